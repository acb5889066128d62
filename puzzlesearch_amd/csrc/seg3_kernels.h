// Kernels of one disk-segment step of the three-bit BFS,
// DB_Opt3BitBFS_Solver::Expand(segment) (Common/DiskBasedOptThreeBitBFS.cpp:54-144).
//
// HBM layout of one segment (2^segment_bits states, at least one word):
//   next   u64 bit array: the reference's IndexedBitArray NextArray; bit i of
//          word i/64 is state i of the segment
//   cur    u64 bit array of the current frontier, only for puzzles with odd
//          cycles (CurArray)
//   dirty  u8 per region of 2^16 states: set by every mark in the region, so
//          the scan passes skip untouched regions (the role of
//          IndexedBitArray's occupancy index, Common/BitArray.h:205-270)
// Both arrays are all zero between steps: the scan clears next, and cur is
// cleared word by word from the current frontier list.
#pragma once

#include "bfs_kernels.h"

namespace psk {

constexpr int S3_REGION_BITS = 16;                               // states per dirty region
constexpr uint32_t S3_REGION_WORDS = 1u << (S3_REGION_BITS - 6);  // u64 words per region (1024)
constexpr uint32_t S3_CROSS_PER_BLOCK = 4096;                    // parents per block of the cross pass

// counters of a step: [0] in-segment children, [1] ids outside the segment,
// [2] the layer count, [3] words scanned, [4] cross-segment children, [5]
// children outside [0, nseg), [6] cross children written, [7] frontier chunks
// that FrontierCompression::Encode would take as a bitmap but are not sorted,
// [8] descents in the old frontier (binned marks need it ascending)
constexpr int S3_NCOUNTERS = 9;

struct S3MarkArgs {
    psd::PuzzleParams prm;
    const uint32_t* idx;  // current frontier, in-segment ids
    uint64_t n;
    uint64_t base;  // segment << seg_bits
    uint32_t segment;
    int seg_bits;
    uint32_t* next;  // the next array as u32 words
    uint8_t* dirty;
    unsigned long long* cur;  // NULL without odd cycles
    unsigned long long* counters;
};

struct S3ScanArgs {
    unsigned long long* next;
    const unsigned long long* cur;  // NULL without odd cycles
    uint8_t* dirty;
    uint64_t nwords;
    unsigned long long* cnt_out;        // per region: states of the new layer
    unsigned long long* counters;
    const unsigned long long* out_off;  // per region exclusive offsets (write pass)
    uint32_t* out_idx;
};

struct S3CrossArgs {
    psd::PuzzleParams prm;
    const uint32_t* idx;  // the new frontier
    uint64_t n;
    uint64_t base;
    uint32_t segment;
    uint32_t nseg;
    int seg_bits;
    unsigned long long* slots;  // per block: (label * nseg + seg) << 32 | in-segment id, compacted
    unsigned long long* counters;
    unsigned long long* block_count;  // per block: cross children written
};

// One OR per next word: consecutive children of a parent in the same 32-bit
// word are merged before the atomic.
struct S3MarkSink {
    uint32_t* next;
    uint8_t* dirty;
    uint64_t segment, in_mask;
    int seg_bits;
    uint64_t pend_w;
    uint32_t pend_m;
    unsigned long long marks;
    __device__ __forceinline__ void flush() {
        if (!pend_m) return;
        gor(&next[pend_w], pend_m);
        dirty[pend_w >> (S3_REGION_BITS - 5)] = 1;  // a plain store: no read to wait for
        pend_m = 0;
    }
    __device__ __forceinline__ void operator()(int, uint64_t child, uint32_t) {
        if ((child >> seg_bits) != segment) return;
        const uint64_t c = child & in_mask;
        ++marks;
        const uint64_t w = c >> 5;
        const uint32_t m = 1u << (c & 31);
        if (w == pend_w) {
            pend_m |= m;
            return;
        }
        flush();
        pend_w = w;
        pend_m = m;
    }
};

// In-segment children of the current frontier set next (and the frontier
// sets cur with odd cycles), Common/DiskBasedOptThreeBitBFS.cpp:78-99.
template <class PZ, int TPB>
__global__ __launch_bounds__(TPB) void k_s3_mark_in(S3MarkArgs a) {
    const uint64_t in_mask = (1ull << a.seg_bits) - 1ull;
    S3MarkSink sink{a.next, a.dirty, a.segment, in_mask, a.seg_bits, ~0ull, 0u, 0ull};
    unsigned long long bad = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * TPB + threadIdx.x; i < a.n; i += (uint64_t)gridDim.x * TPB) {
        const uint32_t id = a.idx[i];
        if ((uint64_t)id > in_mask) {  // not an id of this segment: reported after the step
            ++bad;
            continue;
        }
        if (a.cur) atomicOr(&a.cur[id >> 6], 1ull << (id & 63));
        if constexpr (HasCrossSlots<PZ>::value)  // skip the moves that certainly leave the segment
            PZ::expand(a.base | id, 0u, a.prm, sink, PZ::in_slots(a.base | id, a.seg_bits));
        else
            PZ::expand(a.base | id, 0u, a.prm, sink);
        sink.flush();
        sink.pend_w = ~0ull;
    }
    block_add3<TPB>(sink.marks, bad, 0ull, &a.counters[0], &a.counters[1], nullptr);
}

// incoming cross-segment marks: NextArray.Set(idx) (:64-76)
template <int TPB>
__global__ __launch_bounds__(TPB) void k_s3_mark_list(const uint32_t* __restrict__ idx, uint64_t n, uint64_t in_mask, uint32_t* next,
                               uint8_t* dirty, unsigned long long* bad_count) {
    unsigned long long bad = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t id = idx[i];
        if ((uint64_t)id > in_mask) {
            ++bad;
            continue;
        }
        gor(&next[id >> 5], 1u << (id & 31));
        dirty[id >> S3_REGION_BITS] = 1;
    }
    bad = wave_sum(bad);
    if (bad && threadIdx.x % 64 == 0) atomicAdd(bad_count, bad);
}

// the old frontier: NextArray.Clear(idx) (:101-111); ids outside the segment
// are counted by the mark pass of the frontier that wrote them
template <int TPB>
__global__ __launch_bounds__(TPB) void k_s3_clear_list(const uint32_t* __restrict__ idx, uint64_t n, uint64_t in_mask, uint32_t* next) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t id = idx[i];
        if ((uint64_t)id <= in_mask)
            __hip_atomic_fetch_and(&next[id >> 5], ~(1u << (id & 31)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// CurArray.Clear(): the words holding the current frontier's bits
template <int TPB>
__global__ __launch_bounds__(TPB) void k_s3_clear_cur(const uint32_t* __restrict__ idx, uint64_t n, uint64_t in_mask,
                               unsigned long long* cur) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t id = idx[i];
        if ((uint64_t)id <= in_mask) cur[id >> 6] = 0ull;
    }
}

// thread t of a region's block owns words 4t .. 4t+3 of the region
__device__ __forceinline__ void s3_load4(const S3ScanArgs& a, uint64_t w0, uint64_t rw, uint64_t (&nx)[4],
                                         uint64_t (&live)[4]) {
    const uint64_t j0 = (uint64_t)threadIdx.x * 4;
    if (rw == S3_REGION_WORDS) {
        const ulonglong2* n2 = reinterpret_cast<const ulonglong2*>(a.next + w0 + j0);
        const ulonglong2 x0 = n2[0], x1 = n2[1];
        nx[0] = x0.x;
        nx[1] = x0.y;
        nx[2] = x1.x;
        nx[3] = x1.y;
        if (a.cur) {
            const ulonglong2* c2 = reinterpret_cast<const ulonglong2*>(a.cur + w0 + j0);
            const ulonglong2 y0 = c2[0], y1 = c2[1];
            live[0] = nx[0] & ~y0.x;
            live[1] = nx[1] & ~y0.y;
            live[2] = nx[2] & ~y1.x;
            live[3] = nx[3] & ~y1.y;
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) live[e] = nx[e];
        }
        return;
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const uint64_t j = j0 + e;
        nx[e] = (j < rw) ? a.next[w0 + j] : 0ull;
        live[e] = (a.cur && j < rw) ? nx[e] & ~a.cur[w0 + j] : nx[e];
    }
}

// pass 1 of ScanBitsAndClearWithExcl: per region, the states of the new layer
template <int TPB>
__global__ __launch_bounds__(TPB) void k_s3_count(S3ScanArgs a) {
    static_assert(TPB * 4 == S3_REGION_WORDS, "a region is one block load");
    const uint64_t nregions = (a.nwords + S3_REGION_WORDS - 1) / S3_REGION_WORDS;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    __shared__ uint32_t red[2][TPB / 64];
    unsigned long long total = 0, scanned = 0;  // thread 0
    uint32_t par = 0;
    for (uint64_t r = blockIdx.x; r < nregions; r += gridDim.x, par ^= 1u) {
        const uint64_t w0 = r * S3_REGION_WORDS;
        const uint64_t rw = (a.nwords - w0 < S3_REGION_WORDS) ? a.nwords - w0 : S3_REGION_WORDS;
        uint32_t c = 0;
        const bool dirty = a.dirty[r] != 0;
        if (dirty) {
            uint64_t nx[4], live[4];
            s3_load4(a, w0, rw, nx, live);
#pragma unroll
            for (int e = 0; e < 4; ++e) c += (uint32_t)__popcll(live[e]);
        }
        c = wave_sum(c);
        if (lane == 0) red[par][wave] = c;
        __syncthreads();
        // red[par] is written again two regions later, after the next barrier,
        // which thread 0 reaches only once it has read it here
        if (threadIdx.x == 0) {
            uint32_t s = 0;
#pragma unroll
            for (int w = 0; w < TPB / 64; ++w) s += red[par][w];
            a.cnt_out[r] = s;
            total += s;
            if (dirty) scanned += rw;
        }
    }
    block_add3<TPB>(threadIdx.x == 0 ? total : 0ull, threadIdx.x == 0 ? scanned : 0ull, 0ull, &a.counters[2],
                    &a.counters[3], nullptr);
}

// pass 2: the new frontier's ids in index order, and the clear of next and of
// the region's dirty byte
template <int TPB>
__global__ __launch_bounds__(TPB) void k_s3_write(S3ScanArgs a) {
    constexpr int NW = TPB / 64;
    const uint64_t nregions = (a.nwords + S3_REGION_WORDS - 1) / S3_REGION_WORDS;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    __shared__ uint32_t tot[2][NW];
    uint32_t par = 0;
    for (uint64_t r = blockIdx.x; r < nregions; r += gridDim.x) {
        if (!a.dirty[r]) continue;  // the same value for the whole block: it is cleared after the barrier below
        const uint64_t w0 = r * S3_REGION_WORDS;
        const uint64_t rw = (a.nwords - w0 < S3_REGION_WORDS) ? a.nwords - w0 : S3_REGION_WORDS;
        uint64_t nx[4], live[4];
        s3_load4(a, w0, rw, nx, live);
        uint32_t c = 0;
#pragma unroll
        for (int e = 0; e < 4; ++e) c += (uint32_t)__popcll(live[e]);
        const uint32_t incl = wave_incl_scan(c, lane);
        if (lane == 63) tot[par][wave] = incl;
        __syncthreads();
        // tot[par] is written again two regions later, after the next
        // region's barrier, which every thread reaches only once done here
        uint32_t before = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) before += (w < wave) ? tot[par][w] : 0u;
        uint64_t pos = a.out_off[r] + before + incl - c;
        const uint64_t j0 = w0 + (uint64_t)threadIdx.x * 4;
        // one entry per iteration: a word is fetched when the last one is used up
        uint32_t nz = (live[0] ? 1u : 0u) | (live[1] ? 2u : 0u) | (live[2] ? 4u : 0u) | (live[3] ? 8u : 0u);
        uint64_t v = 0, wid = 0;
        for (;;) {
            if (!v) {
                if (!nz) break;
                const int e = __ffs(nz) - 1;
                nz &= nz - 1u;
                v = (e == 0) ? live[0] : (e == 1) ? live[1] : (e == 2) ? live[2] : live[3];
                wid = (j0 + (uint64_t)e) * 64;
            }
            const int b = __ffsll((unsigned long long)v) - 1;
            v &= v - 1;
            a.out_idx[pos++] = (uint32_t)(wid + (uint64_t)b);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e)
            if (nx[e]) a.next[j0 + e] = 0ull;
        if (threadIdx.x == 0) a.dirty[r] = 0;  // every thread read it before the barrier above
        par ^= 1u;
    }
}

// children of the new frontier that leave the segment, in slot order
struct S3CrossSink {
    uint64_t segment, nseg, in_mask;
    int seg_bits;
    unsigned long long bad;
    uint32_t valid;             // bit k: slot k holds a cross child
    unsigned long long v[16];   // by slot; PZ::SLOTS <= 16
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

// Block b expands parents [b*4096, (b+1)*4096) with the cross-segment hint
// and writes their cross children, compacted in (parent, slot) order, to the
// front of its own area slots[b*4096*SLOTS ...]; block_count[b] is how many.
template <class PZ, int TPB>
__global__ __launch_bounds__(TPB) void k_s3_cross(S3CrossArgs a) {
    static_assert(PZ::SLOTS <= 16, "cross sink holds 16 slots");
    constexpr int NW = TPB / 64;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    S3CrossSink sink{a.segment, a.nseg, (1ull << a.seg_bits) - 1ull, a.seg_bits, 0ull, 0u, {}};
    const uint64_t p0 = (uint64_t)blockIdx.x * S3_CROSS_PER_BLOCK;
    const uint64_t p1 = (p0 + S3_CROSS_PER_BLOCK < a.n) ? p0 + S3_CROSS_PER_BLOCK : a.n;
    unsigned long long* out = a.slots + p0 * PZ::SLOTS;
    __shared__ uint32_t tot[2][NW];
    uint32_t par = 0, running = 0;
    unsigned long long nx = 0;
    for (uint64_t k0 = p0; k0 < p1; k0 += TPB, par ^= 1u) {
        const uint64_t i = k0 + threadIdx.x;
        sink.valid = 0;
        if (i < p1) PZ::expand(a.base | a.idx[i], 0u, a.prm, sink);
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
    block_add3<TPB>(nx, sink.bad, nx, &a.counters[4], &a.counters[5], &a.counters[6]);
    if (threadIdx.x == 0) a.block_count[blockIdx.x] = running;
}

// ---------------------------------------------------------------- binned marks

// Without the NextArray in HBM (puzzles expanded in full, at most 4 moves):
// the marks are binned per region of 2^rb states by the bin passes of the
// frontier-search step (k_fs_bin, labels dropped), and one persistent
// workgroup per CU builds each region's NextArray bits in LDS from them,
// clears the old frontier's bits (its ids are sorted, so each region's are a
// range found by binary search), excludes the current frontier's bits (odd
// cycles, read from CurArray) and counts (WRITE = false) or writes the new
// layer's ids in order (WRITE = true).  Thread t owns bit words [8t, 8t+8).
// The write pass needs S3_STAGE_IDS * 4 bytes of dynamic LDS.
struct S3RegionArgs {
    uint32_t nregions;
    int rb;
    const uint32_t* marks;             // (state in region << 2), region by region
    const unsigned long long* rstart;  // nregions + 1
    const uint32_t* old;               // the old frontier, ascending in-segment ids
    const unsigned long long* ostart;  // nregions + 1: each region's range of old
    const unsigned long long* cur;     // CurArray (odd cycles), else NULL
    unsigned long long* cnt_out;       // count pass: per region
    unsigned long long* counters;      // count pass: [2] += count, [3] += words scanned
    const unsigned long long* out_off; // write pass
    uint32_t* out_idx;
};

constexpr int S3_BIN_TPB = 1024;
constexpr int S3_BIN_WPT = 8;  // LDS bit words per thread: 2^18 states per region
// the write pass stages up to this many ids of a region in (dynamic) LDS and
// stores them with consecutive lanes on consecutive ids; a region with more
// ids is written straight from the owner threads
constexpr uint32_t S3_STAGE_IDS = 11264;  // 44 KB: with the 32 KB of bits two workgroups fit a CU

// ostart[r] = first index of `ids` (ascending) at or above r << rb, r = 0..nregions
template <int TPB>
__global__ __launch_bounds__(TPB) void k_s3_range_starts(const uint32_t* ids, uint64_t n, int rb, uint32_t nregions,
                                                         unsigned long long* ostart) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r > nregions) return;
    const uint64_t key = (uint64_t)r << rb;
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) / 2;
        if ((uint64_t)ids[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    ostart[r] = lo;
}

// ids above in_mask (counted in *bad) and, with `desc`, places where an id
// is below its predecessor (counted in *desc)
template <int TPB>
__global__ __launch_bounds__(TPB) void k_s3_check_list(const uint32_t* ids, uint64_t n, uint64_t in_mask,
                                                       unsigned long long* bad, unsigned long long* desc) {
    unsigned long long nb = 0, nd = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t id = ids[i];
        if ((uint64_t)id > in_mask) ++nb;
        if (desc && i > 0 && id < ids[i - 1]) ++nd;
    }
    nb = wave_sum(nb);
    nd = wave_sum(nd);
    if (threadIdx.x % 64 == 0) {
        if (nb && bad) atomicAdd(bad, nb);
        if (nd) atomicAdd(desc, nd);
    }
}

template <int TPB, bool WRITE>
__global__ __launch_bounds__(TPB) void k_s3_region(S3RegionArgs a) {
    constexpr int NW = TPB / 64;
    constexpr int WPT = S3_BIN_WPT;
    static_assert(TPB * WPT * 32 == (1 << FS_BIN_BITS), "a region's bit words fill the workgroup");
    __shared__ uint32_t bits[TPB * WPT];
    __shared__ uint32_t s_w[NW];
    extern __shared__ uint32_t stage[];  // write pass: S3_STAGE_IDS ids
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const uint32_t rwords = (1u << a.rb) / 32u;  // bit words of the region (<= TPB * WPT)
    const uint32_t t8 = threadIdx.x * WPT;
    for (uint32_t x = threadIdx.x; x < TPB * WPT; x += TPB) bits[x] = 0u;
    unsigned long long total = 0, scanned = 0;  // thread 0
    for (uint32_t r = blockIdx.x; r < a.nregions; r += gridDim.x) {
        const uint64_t m0 = a.rstart[r], m1 = a.rstart[r + 1];
        if (m0 == m1) {  // no mark: nothing in the new layer here
            if (!WRITE && threadIdx.x == 0) a.cnt_out[r] = 0;
            continue;
        }
        __syncthreads();  // the bits are clear (start, or cleared by their owners)
        for (uint64_t j = m0 + threadIdx.x; j < m1; j += TPB) {
            const uint32_t st = a.marks[j] >> 2;
            lor(&bits[st >> 5], 1u << (st & 31u));
        }
        __syncthreads();
        const uint32_t rmask = (1u << a.rb) - 1u;
        for (uint64_t j = a.ostart[r] + threadIdx.x; j < a.ostart[r + 1]; j += TPB) {
            const uint32_t st = a.old[j] & rmask;
            atomicAnd(&bits[st >> 5], ~(1u << (st & 31u)));
        }
        __syncthreads();
        uint32_t live[WPT];
        uint32_t c = 0;
        const uint64_t g0 = (uint64_t)r * rwords + t8;  // this thread's first bit word in the segment (32-bit words)
#pragma unroll
        for (int k = 0; k < WPT; ++k) {
            live[k] = (t8 + (uint32_t)k < rwords) ? bits[t8 + k] : 0u;
            if (a.cur && live[k]) live[k] &= ~(uint32_t)(a.cur[(g0 + (uint64_t)k) >> 1] >> (32 * ((g0 + (uint64_t)k) & 1)));
            c += (uint32_t)__popc(live[k]);
        }
        if (!WRITE) {
            const uint32_t wc = wave_sum(c);
            if (lane == 0) s_w[wave] = wc;
            __syncthreads();
            if (threadIdx.x == 0) {
                uint32_t sum = 0;
                for (int x = 0; x < NW; ++x) sum += s_w[x];
                a.cnt_out[r] = sum;
                total += sum;
                scanned += rwords / 2;  // in 64-bit words
            }
        } else {
            const uint32_t incl = wave_incl_scan(c, lane);
            if (lane == 63) s_w[wave] = incl;
            __syncthreads();
            const uint32_t wt = lane < NW ? s_w[lane] : 0u;
            const uint32_t wi = wave_incl_scan(wt, lane);
            const uint64_t r0 = a.out_off[r], nr = a.out_off[r + 1] - r0;
            const bool staged = nr <= S3_STAGE_IDS;  // the same for the whole workgroup
            uint32_t* dst = staged ? stage : a.out_idx + r0;
            uint64_t pos = __shfl(wi - wt, wave, 64) + incl - c;
            // one entry per iteration: a word is fetched when the last one is used up
            uint32_t nz = 0;
#pragma unroll
            for (int k = 0; k < WPT; ++k) nz |= live[k] ? (1u << k) : 0u;
            uint32_t v = 0, base = 0;
            for (;;) {
                if (!v) {
                    if (!nz) break;
                    const int k = __ffs(nz) - 1;
                    nz &= nz - 1u;
                    uint32_t x = 0;
#pragma unroll
                    for (int q = 0; q < WPT; ++q) x = (q == k) ? live[q] : x;
                    v = x;
                    base = (uint32_t)((g0 + (uint64_t)k) * 32u);
                }
                const int b = __ffs(v) - 1;
                v &= v - 1u;
                dst[pos++] = base + (uint32_t)b;
            }
            if (staged) {
                __syncthreads();
                for (uint32_t i = threadIdx.x; i < (uint32_t)nr; i += TPB) a.out_idx[r0 + i] = stage[i];
                __syncthreads();  // the stage is free for the next region
            }
        }
        // this thread's words are read only by this thread: clear them for the next region
#pragma unroll
        for (int k = 0; k < WPT; ++k)
            if (t8 + (uint32_t)k < rwords) bits[t8 + k] = 0u;
    }
    if (!WRITE) block_add3<TPB>(total, scanned, 0ull, &a.counters[2], &a.counters[3], nullptr);
}

}  // namespace psk
