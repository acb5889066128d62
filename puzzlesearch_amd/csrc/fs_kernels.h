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

// puzzles that can name the slots whose child may leave the segment
// (PZ::cross_slots(index, seg_bits)); the others expand every slot
template <class PZ, class = void>
struct HasCrossSlots : std::false_type {};
template <class PZ>
struct HasCrossSlots<PZ, std::void_t<decltype(PZ::cross_slots(0ull, 0))>> : std::true_type {};

// One OR per flag word: consecutive children of a parent that share a word
// (the horizontal moves of the sliding puzzles: index -+ 1) are merged before
// the atomic.
struct FsMarkSink {
    uint32_t* flags;
    uint8_t* dirty;
    uint64_t segment, in_mask;
    int seg_bits;
    uint64_t pend_w;
    uint32_t pend_m;
    unsigned long long marks;
    __device__ __forceinline__ void flush() {
        if (!pend_m) return;
        gor(&flags[pend_w], pend_m);
        dirty[pend_w >> (FS_REGION_BITS - 3)] = 1;  // a plain store: no read to wait for
        pend_m = 0;
    }
    __device__ __forceinline__ void operator()(int, uint64_t child, uint32_t label) {
        if ((child >> seg_bits) != segment) return;
        const uint64_t c = child & in_mask;
        const uint32_t m = (1u << label) << ((c & 7) * 4);
        ++marks;
        if ((c >> 3) == pend_w) {
            pend_m |= m;
            return;
        }
        flush();
        pend_w = c >> 3;
        pend_m = m;
    }
};

template <class PZ, int TPB>
__global__ __launch_bounds__(TPB) void k_fs_mark_in(FsMarkArgs a) {
    const uint64_t in_mask = (1ull << a.seg_bits) - 1ull;
    FsMarkSink sink{a.flags, a.dirty, a.segment, in_mask, a.seg_bits, ~0ull, 0u, 0ull};
    unsigned long long bad = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * TPB + threadIdx.x; i < a.n; i += (uint64_t)gridDim.x * TPB) {
        const uint32_t id = a.idx[i];
        if ((uint64_t)id > in_mask) {  // not an id of this segment: reported after the step
            ++bad;
            continue;
        }
        if (a.odd) atomicOr(&a.odd[id >> 6], 1ull << (id & 63));
        if constexpr (HasCrossSlots<PZ>::value)  // skip the moves that certainly leave the segment
            PZ::expand(a.base | id, (uint32_t)a.ops[i], a.prm, sink, PZ::in_slots(a.base | id, a.seg_bits));
        else
            PZ::expand(a.base | id, (uint32_t)a.ops[i], a.prm, sink);
        sink.flush();
        sink.pend_w = ~0ull;
    }
    block_add3<TPB>(sink.marks, bad, 0ull, &a.counters[0], &a.counters[6], nullptr);
}


// pass 1 of ScanBitsAndClear: per region, entries counted and entries written.
// Persistent blocks take regions round-robin; a full region is read as
// FS_REGION_WORDS/4/TPB 16-byte loads per thread, all issued before any is
// used.  The totals go to the counters once per block.
template <int TPB>
__global__ __launch_bounds__(TPB) void k_fs_count(FsScanArgs a) {
    constexpr uint32_t PER = FS_REGION_WORDS / 4 / TPB;
    static_assert(PER * 4 * TPB == FS_REGION_WORDS, "a region is a whole number of block loads");
    const uint64_t nregions = (a.nwords + FS_REGION_WORDS - 1) / FS_REGION_WORDS;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    __shared__ uint32_t red[2][TPB / 64];
    unsigned long long all = 0;       // per thread, over every region of the block
    unsigned long long out_blk = 0, scanned = 0;  // thread 0
    uint32_t par = 0;
    for (uint64_t r = blockIdx.x; r < nregions; r += gridDim.x, par ^= 1u) {
        const uint64_t w0 = r * FS_REGION_WORDS;
        const uint64_t rw = (a.nwords - w0 < FS_REGION_WORDS) ? a.nwords - w0 : FS_REGION_WORDS;
        uint32_t out = 0;
        const bool dirty = a.dirty[r] != 0;
        if (dirty && rw == FS_REGION_WORDS) {
            const uint4* f4 = reinterpret_cast<const uint4*>(a.flags_c + w0);
            uint4 v[PER];
#pragma unroll
            for (uint32_t k = 0; k < PER; ++k) v[k] = f4[k * TPB + threadIdx.x];
#pragma unroll
            for (uint32_t k = 0; k < PER; ++k) {
                const uint32_t ws[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    if (!ws[e]) continue;
                    const uint32_t live = fs_live(ws[e], a.odd, w0 + ((uint64_t)k * TPB + threadIdx.x) * 4 + e);
                    all += __popc(live);
                    out += __popc(live & ~nib_full(ws[e], a.ops));
                }
            }
        } else if (dirty) {  // the whole segment is smaller than a region
            for (uint64_t j = threadIdx.x; j < rw; j += TPB) {
                const uint32_t w = a.flags_c[w0 + j];
                if (!w) continue;
                const uint32_t live = fs_live(w, a.odd, w0 + j);
                all += __popc(live);
                out += __popc(live & ~nib_full(w, a.ops));
            }
        }
        out = wave_sum(out);
        if (lane == 0) red[par][wave] = out;
        __syncthreads();
        // red[par] is written again two regions later, after the next barrier,
        // which thread 0 reaches only once it has read it here
        if (threadIdx.x == 0) {
            uint32_t so = 0;
#pragma unroll
            for (int w = 0; w < TPB / 64; ++w) so += red[par][w];
            a.cnt_out[r] = so;
            out_blk += so;
            if (dirty) scanned += rw;
        }
    }
    block_add3<TPB>(all, threadIdx.x == 0 ? out_blk : 0ull, threadIdx.x == 0 ? scanned : 0ull, &a.counters[1],
                    &a.counters[2], &a.counters[5]);
}

// pass 2: ordered write of (id, op bits) for entries with an unused move, and
// the clear of the region (flags words and dirty byte).  Persistent blocks
// take regions round-robin.  In a full region thread t of load k owns words
// 4*(k*TPB + t) .. +3, so (load, thread, word) order is index order: the
// block scans the per-load counts of CH loads at a time after one barrier
// and writes every entry at its sorted position.
template <int TPB>
__global__ __launch_bounds__(TPB) void k_fs_write(FsScanArgs a) {
    constexpr uint32_t PER = FS_REGION_WORDS / 4 / TPB;
    constexpr uint32_t CH = PER < 4 ? PER : 4;
    static_assert(PER % CH == 0, "a region is a whole number of chunks");
    constexpr int NW = TPB / 64;
    const uint64_t nregions = (a.nwords + FS_REGION_WORDS - 1) / FS_REGION_WORDS;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    __shared__ uint32_t tot[2][CH][NW];
    uint32_t par = 0;
    for (uint64_t r = blockIdx.x; r < nregions; r += gridDim.x) {
        if (!a.dirty[r]) continue;  // the same value for the whole block: it is cleared after the barrier below
        const uint64_t w0 = r * FS_REGION_WORDS;
        const uint64_t rw = (a.nwords - w0 < FS_REGION_WORDS) ? a.nwords - w0 : FS_REGION_WORDS;
        uint64_t running = a.out_off[r];
        if (rw == FS_REGION_WORDS) {
            uint4* f4 = reinterpret_cast<uint4*>(a.flags + w0);
            // the region in PER/CH chunks of CH loads per thread, one barrier each
            for (uint32_t k0 = 0; k0 < PER; k0 += CH, par ^= 1u) {
                uint4 v[CH];
#pragma unroll
                for (uint32_t k = 0; k < CH; ++k) v[k] = f4[(k0 + k) * TPB + threadIdx.x];
                uint32_t c[CH], incl[CH], keep[CH][4];
#pragma unroll
                for (uint32_t k = 0; k < CH; ++k) {
                    const uint32_t ws[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
                    const uint64_t j0 = w0 + ((uint64_t)(k0 + k) * TPB + threadIdx.x) * 4;
                    c[k] = 0;
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        keep[k][e] = ws[e] ? fs_live(ws[e], a.odd, j0 + e) & ~nib_full(ws[e], a.ops) : 0u;
                        c[k] += __popc(keep[k][e]);
                    }
                    incl[k] = wave_incl_scan(c[k], lane);
                    if (lane == 63) tot[par][k][wave] = incl[k];
                }
                __syncthreads();
                // tot[par] is written again two chunks later, after the next
                // chunk's barrier, which every thread reaches only once done here
#pragma unroll
                for (uint32_t k = 0; k < CH; ++k) {
                    uint32_t before = 0, total = 0;
#pragma unroll
                    for (int w = 0; w < NW; ++w) {
                        const uint32_t t = tot[par][k][w];
                        before += (w < wave) ? t : 0u;
                        total += t;
                    }
                    const uint32_t ws[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
                    const uint64_t j0 = w0 + ((uint64_t)(k0 + k) * TPB + threadIdx.x) * 4;
                    uint64_t pos = running + before + incl[k] - c[k];
                    for (int e = 0; e < 4; ++e) {
                        uint32_t kb = keep[k][e];
                        while (kb) {
                            const int b = __ffs(kb) - 1;  // bit 4k of nibble k
                            kb &= kb - 1;
                            a.out_idx[pos] = (uint32_t)((j0 + e) * 8 + (uint64_t)(b >> 2));
                            a.out_ops[pos] = (uint8_t)((ws[e] >> b) & 15u);
                            ++pos;
                        }
                    }
                    if (ws[0] | ws[1] | ws[2] | ws[3]) f4[(k0 + k) * TPB + threadIdx.x] = make_uint4(0u, 0u, 0u, 0u);
                    running += total;
                }
            }
            if (threadIdx.x == 0) a.dirty[r] = 0;  // every thread read it before the first barrier above
            continue;
        }
        // the whole segment is smaller than a region (nregions == 1): thread t
        // of iteration `it` owns words it*4*TPB + 4t .. +3, one block scan per iteration
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
            if (lane == 63) tot[0][0][wave] = incl;
            __syncthreads();
            uint32_t before = 0, total = 0;
            for (int k = 0; k < NW; ++k) {
                const uint32_t t = tot[0][0][k];
                before += (k < wave) ? t : 0u;
                total += t;
            }
            uint64_t pos = running + before + incl - c;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                uint32_t kb = keep[e];
                while (kb) {
                    const int b = __ffs(kb) - 1;
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
}


// children of the written entries that leave the segment, in slot order
struct CrossSink {
    uint64_t segment, nseg, in_mask;
    int seg_bits;
    unsigned long long bad;
    uint32_t valid;            // bit k: slot k holds a cross child
    unsigned long long v[8];   // by slot; PZ::SLOTS <= 8
    __device__ __forceinline__ void operator()(int slot, uint64_t child, uint32_t label) {
        const uint64_t s = child >> seg_bits;
        if (s == segment) return;
        if (s >= nseg) {
            ++bad;
            return;
        }
        v[slot] = (((uint64_t)label * nseg + s) << 32) | (child & in_mask);
        valid |= 1u << slot;
    }
};

constexpr uint32_t FS_CROSS_PER_BLOCK = 4096;  // parents per block of the cross pass


// Block b expands parents [b*4096, (b+1)*4096) and writes their cross
// children, compacted in (parent, slot) order, to the front of its own area
// slots[b*4096*SLOTS ...]; block_count[b] is how many it wrote, so the
// compaction copies exactly sum(block_count) entries.  counters[7] receives
// the same total for the host to compare with the children counted in
// counters[3].  For puzzles that name the slots that may leave the segment
// the block first queues, in parent order, only the parents with such a
// slot, so the lanes of a wave expand parents that all have work.
template <class PZ, int TPB>
__global__ __launch_bounds__(TPB) void k_fs_cross(FsCrossArgs a) {
    static_assert(PZ::SLOTS <= 8, "cross sink holds 8 slots");
    constexpr int NW = TPB / 64;
    constexpr bool QUEUE = HasCrossSlots<PZ>::value;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    CrossSink sink{a.segment, a.nseg, (1ull << a.seg_bits) - 1ull, a.seg_bits, 0ull, 0u, {}};
    const uint64_t p0 = (uint64_t)blockIdx.x * FS_CROSS_PER_BLOCK;
    const uint64_t p1 = (p0 + FS_CROSS_PER_BLOCK < a.n) ? p0 + FS_CROSS_PER_BLOCK : a.n;
    unsigned long long* out = a.slots + p0 * PZ::SLOTS;
    __shared__ uint32_t tot[2][NW];
    __shared__ uint16_t q[QUEUE ? FS_CROSS_PER_BLOCK : 1];
    uint32_t par = 0;
    uint32_t nq = (uint32_t)(p1 - p0);
    if constexpr (QUEUE) {
        uint32_t qn = 0;
        for (uint64_t b0 = p0; b0 < p1; b0 += TPB, par ^= 1u) {
            const uint64_t i = b0 + threadIdx.x;
            bool need = false;
            if (i < p1) {
                const uint64_t parent = a.base | a.idx[i];
                need = (PZ::cross_slots(parent, a.seg_bits) & PZ::live_slots(parent, (uint32_t)a.ops[i])) != 0u;
            }
            const unsigned long long m = __ballot(need);
            if (lane == 0) tot[par][wave] = (uint32_t)__popcll(m);
            __syncthreads();
            uint32_t before = 0, total = 0;
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                const uint32_t t = tot[par][w];
                before += (w < wave) ? t : 0u;
                total += t;
            }
            if (need) q[qn + before + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = (uint16_t)(i - p0);
            qn += total;
        }
        __syncthreads();  // the queue is complete
        nq = qn;
    }
    uint32_t running = 0;
    unsigned long long nx = 0;
    for (uint32_t k0 = 0; k0 < nq; k0 += TPB, par ^= 1u) {
        const uint32_t k = k0 + threadIdx.x;
        sink.valid = 0;
        if (k < nq) {
            const uint64_t i = p0 + (QUEUE ? (uint64_t)q[k] : (uint64_t)k);
            const uint64_t parent = a.base | a.idx[i];
            if constexpr (QUEUE)
                PZ::expand(parent, (uint32_t)a.ops[i], a.prm, sink, PZ::cross_slots(parent, a.seg_bits));
            else
                PZ::expand(parent, (uint32_t)a.ops[i], a.prm, sink);
        }
        const uint32_t c = (uint32_t)__popc(sink.valid);
        nx += c;
        const uint32_t incl = wave_incl_scan(c, lane);
        if (lane == 63) tot[par][wave] = incl;
        __syncthreads();
        // tot[par] is written again two iterations later, after the next
        // iteration's barrier, which every thread reaches only once done here
        uint32_t before = 0, total = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const uint32_t t = tot[par][w];
            before += (w < wave) ? t : 0u;
            total += t;
        }
        uint32_t pos = running + before + incl - c;
#pragma unroll
        for (int kk = 0; kk < PZ::SLOTS; ++kk)
            if (sink.valid & (1u << kk)) out[pos++] = sink.v[kk];
        running += total;
    }
    block_add3<TPB>(nx, sink.bad, nx, &a.counters[3], &a.counters[4], &a.counters[7]);
    if (threadIdx.x == 0) a.block_count[blockIdx.x] = running;
}

// Ordered compaction of the cross children: block b copies the block_count[b]
// entries at the front of its area (slots_per_block entries from
// slots[b*slots_per_block]) to out[block_off[b] ...].
template <int TPB>
__global__ __launch_bounds__(TPB) void k_fs_compact(const unsigned long long* __restrict__ slots,
                                                    uint32_t slots_per_block,
                                                    const unsigned long long* __restrict__ block_count,
                                                    const unsigned long long* __restrict__ block_off,
                                                    unsigned long long* __restrict__ out) {
    const unsigned long long* src = slots + (uint64_t)blockIdx.x * slots_per_block;
    unsigned long long* dst = out + block_off[blockIdx.x];
    const uint64_t n = block_count[blockIdx.x];
    for (uint64_t i = threadIdx.x; i < n; i += TPB) dst[i] = src[i];
}


}  // namespace psk
