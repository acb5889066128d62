// Kernels of one disk-segment step of the frontier search with 4-bit op flags,
// DB_OptFS_Solver<4>::Expand(segment) (Common/DiskBasedOptFrontierSearch.cpp:53-132).
//
// The op flags of a segment (2^segment_bits states, 4 bits each: the
// IndexedArray<4> of the reference, Common/BitArray.h:283-314) are never held
// in HBM.  The step bins its marks instead:
//   count  each workgroup takes a contiguous share of the frontier and of the
//          incoming cross-segment marks and counts their in-segment marks per
//          region (regions of 2^FS_BIN_BITS states) in LDS, one row of a
//          (workgroup, region) table;
//   place  after a prefix sum down every region's column and over the regions,
//          each workgroup writes the same marks to its own part of their
//          region's range of a mark list (state in region << 2 | move), taking
//          positions from LDS cursors;
//   region one workgroup per region ORs the region's marks into an LDS copy
//          of its flags (128 KB, with a bitmap of the words holding a mark)
//          and counts the entries, then, after a prefix sum of the
//          per-region output counts, rebuilds them and writes the region's
//          (id, op bits) in id order.
// For puzzles with odd cycles the current frontier's bits (`odd`, a u64 bit
// array of the segment) exclude its entries, as the reference's
// FrontierArray does.
#pragma once

#include "bfs_kernels.h"

namespace psk {

constexpr int FS_BIN_BITS = 18;   // states per region: 2^18, flags of 128 KB
constexpr uint64_t FS_NONE = ~0ull;  // empty cross-segment slot
constexpr int FS_REGION_TPB = 1024;  // thread t of a region workgroup owns flag words [32t, 32t + 32)
static_assert((1 << FS_BIN_BITS) / 8 <= 32 * FS_REGION_TPB, "a region's flag words fit its workgroup");
constexpr uint32_t FS_MAX_REGIONS = 1u << (32 - FS_BIN_BITS);  // LDS cursors of one workgroup: 64 KB
constexpr int FS_BIN_TPB = 1024;
constexpr uint32_t FS_MAX_BIN_BLOCKS = 512;  // rows of the (workgroup, region) table
constexpr int FS_MK = 8;                      // marks a thread of the region pass loads at once
constexpr int FS_COL_WAVES = 16;              // column pass over the table: waves per workgroup
constexpr int FS_COL_ROWS = 32;               // and rows per wave
static_assert(FS_COL_WAVES * FS_COL_ROWS == (int)FS_MAX_BIN_BLOCKS, "column pass covers every row");

struct FsBinArgs {
    psd::PuzzleParams prm;
    const uint32_t* idx;  // frontier, in-segment ids
    const uint8_t* ops;   // their op bits (NULL: none)
    uint64_t n;
    uint64_t base;  // segment << seg_bits
    uint32_t segment;
    int seg_bits;
    int rb;                       // log2(states per region) = min(FS_BIN_BITS, seg_bits)
    uint32_t nregions;            // at most FS_MAX_REGIONS
    uint64_t per, perx;           // parents and cross marks per workgroup
    const uint32_t* xidx;         // incoming cross-segment marks, op by op
    const uint64_t* xstarts;      // ops + 1 entries (device)
    uint64_t nx;
    int nops;
    unsigned long long* odd;      // frontier bits (odd-cycle puzzles; count pass), else NULL
    uint32_t* table;              // gridDim.x rows of nregions: count pass, the workgroup's marks per region;
                                  // place pass, its first position relative to the region's start
    const unsigned long long* rstart;  // place pass: nregions + 1 region starts in the mark list
    uint32_t* marks;              // place pass: (state in region << 2) | move
    unsigned long long* children_out;  // count pass: += in-segment children of the frontier
    unsigned long long* bad_out;       // count pass: += frontier ids outside the segment
    uint32_t label_mask;               // 3: the move labels of the 4-bit op flags; 0: marks without labels
};

struct FsRegionArgs {
    uint32_t nregions;
    const uint32_t* marks;
    const unsigned long long* rstart;  // nregions + 1
    const unsigned long long* odd;     // NULL without odd cycles
    int rb;
    int ops;
    unsigned long long* cnt_out;        // count pass: entries written per region
    uint32_t* cnt_all;                  // count pass: entries per region
    const unsigned long long* out_off;  // write pass: exclusive offsets
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

// The in-segment children of one parent, by slot.
template <int SLOTS>
struct BinSink {
    uint64_t segment, in_mask;
    int seg_bits;
    uint32_t valid;  // bit k: slot k holds an in-segment child
    uint32_t v[SLOTS];
    uint32_t lab[SLOTS];
    __device__ __forceinline__ void operator()(int slot, uint64_t child, uint32_t label) {
        if ((child >> seg_bits) != segment) return;
        v[slot] = (uint32_t)(child & in_mask);
        lab[slot] = label;
        valid |= 1u << slot;
    }
};

// Count (PLACE = false) or place (PLACE = true) the marks of the frontier's
// in-segment children and of the incoming cross-segment marks.  Workgroup b
// takes parents [b*per, (b+1)*per) and cross marks [b*perx, (b+1)*perx); the
// place pass sees the same marks as the count pass, so each workgroup fills
// exactly the positions its table row reserved.
template <class PZ, int TPB, bool PLACE>
__global__ __launch_bounds__(TPB) void k_fs_bin(FsBinArgs a) {
    extern __shared__ uint32_t cur[];  // per region: marks counted, or the next position
    uint32_t* row = a.table + (uint64_t)blockIdx.x * a.nregions;
    for (uint32_t r = threadIdx.x; r < a.nregions; r += TPB)
        cur[r] = PLACE ? (uint32_t)a.rstart[r] + row[r] : 0u;
    __syncthreads();
    const uint64_t in_mask = (1ull << a.seg_bits) - 1ull;
    const uint32_t lmask = (1u << a.rb) - 1u;
    auto put = [&](uint32_t state, uint32_t label) {
        uint32_t* c = &cur[state >> a.rb];
        if constexpr (PLACE) a.marks[atomicAdd(c, 1u)] = ((state & lmask) << 2) | (label & a.label_mask);
        else atomicAdd(c, 1u);
    };
    unsigned long long marks = 0, bad = 0;
    const uint64_t p0 = (uint64_t)blockIdx.x * a.per < a.n ? (uint64_t)blockIdx.x * a.per : a.n;
    const uint64_t p1 = p0 + a.per < a.n ? p0 + a.per : a.n;
    for (uint64_t i = p0 + threadIdx.x; i < p1; i += TPB) {
        BinSink<PZ::SLOTS> sink{a.segment, in_mask, a.seg_bits, 0u, {}, {}};
        const uint32_t id = a.idx[i];
        if ((uint64_t)id > in_mask) {  // not an id of this segment: reported after the step
            if (!PLACE) ++bad;
            continue;
        }
        if (!PLACE && a.odd) atomicOr(&a.odd[id >> 6], 1ull << (id & 63));
        const uint32_t opb = a.ops ? (uint32_t)a.ops[i] : 0u;
        if constexpr (HasCrossSlots<PZ>::value)  // skip the moves that certainly leave the segment
            PZ::expand(a.base | id, opb, a.prm, sink, PZ::in_slots(a.base | id, a.seg_bits));
        else
            PZ::expand(a.base | id, opb, a.prm, sink);
        marks += (unsigned long long)__popc(sink.valid);
#pragma unroll
        for (int k = 0; k < PZ::SLOTS; ++k)
            if ((sink.valid >> k) & 1u) put(sink.v[k], sink.lab[k]);
    }
    const uint64_t x0 = (uint64_t)blockIdx.x * a.perx < a.nx ? (uint64_t)blockIdx.x * a.perx : a.nx;
    const uint64_t x1 = x0 + a.perx < a.nx ? x0 + a.perx : a.nx;
    for (uint64_t j = x0 + threadIdx.x; j < x1; j += TPB) {
        uint32_t op = 0;
        while ((int)op + 1 < a.nops && a.xstarts[op + 1] <= j) ++op;
        put((uint32_t)(a.xidx[j] & in_mask), op);
    }
    if (!PLACE) {
        __syncthreads();
        for (uint32_t r = threadIdx.x; r < a.nregions; r += TPB) row[r] = cur[r];
        block_add3<TPB>(marks, bad, 0ull, a.children_out, a.bad_out, nullptr);
    }
}

// One workgroup per region: its flags are built in LDS from its marks, then
// counted (WRITE = false: all entries, and entries with an unused move, per
// region) or written in id order (WRITE = true).  Workgroups are persistent
// (workgroup b takes regions b, b + grid, ...) and load the next region's mark
// range, output offset and first FS_MK marks per thread while they work on the
// current one.  Beside the flags the workgroup keeps one bit per flag word
// (set when a mark lands in the word); thread t owns flag words
// [32t, 32t + 32), visits only its words with a mark, writes its entries in id
// order after a workgroup scan of the per-thread counts, and clears its words
// again for the next region.
template <int TPB, bool WRITE>
__global__ __launch_bounds__(TPB) void k_fs_region(FsRegionArgs a) {
    extern __shared__ uint32_t fl[];  // rw flag words, then nbm words of the nonzero-word bitmap
    constexpr int NW = TPB / 64;
    static_assert(NW <= 64, "one lane per wave total");
    const uint32_t rw = (1u << a.rb) / 8u;  // flag words of the region, at most 32 * TPB
    const uint32_t nbm = (rw + 31u) / 32u;
    uint32_t* bm = fl + rw;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const uint32_t t32 = threadIdx.x * 32u;
    // per-wave totals; written only after the barrier that opens a region, which
    // every thread reaches once done reading the previous region's totals
    __shared__ uint32_t s_w[NW];
    for (uint32_t x = threadIdx.x; x < rw + nbm; x += TPB) fl[x] = 0u;
    uint32_t c[FS_MK];
    auto load = [&](uint64_t j0, uint64_t end) {
#pragma unroll
        for (int k = 0; k < FS_MK; ++k) {
            const uint64_t j = j0 + (uint64_t)k * TPB + threadIdx.x;
            c[k] = j < end ? a.marks[j] : ~0u;  // a mark is below 2^20
        }
    };
    // a word's entries: nonzero, not in the current frontier, some move unused
    auto keep_of = [&](uint32_t f, uint64_t word) -> uint32_t {
        return fs_live(f, a.odd, word) & ~nib_full(f, a.ops);
    };
    uint32_t r = blockIdx.x;
    uint64_t m0 = 0, m1 = 0, off = 0;
    if (r < a.nregions) {
        m0 = a.rstart[r];
        m1 = a.rstart[r + 1];
        if (WRITE) off = a.out_off[r];
        load(m0, m1);
    }
    for (; r < a.nregions; r += gridDim.x) {
        const uint32_t rn = r + gridDim.x;
        uint64_t n0 = 0, n1 = 0, noff = 0;
        if (rn < a.nregions) {
            n0 = a.rstart[rn];
            n1 = a.rstart[rn + 1];
            if (WRITE) noff = a.out_off[rn];
        }
        if (m0 == m1) {  // no mark: nothing to count or write
            if (!WRITE && threadIdx.x == 0) {
                a.cnt_out[r] = 0;
                a.cnt_all[r] = 0;
            }
            load(n0, n1);
            m0 = n0;
            m1 = n1;
            off = noff;
            continue;
        }
        __syncthreads();  // flags and bitmap are clear (first region, or cleared by their owners)
        for (uint64_t j0 = m0;;) {
#pragma unroll
            for (int k = 0; k < FS_MK; ++k) {
                if (c[k] == ~0u) continue;
                const uint32_t st = c[k] >> 2;
                const uint32_t w = st >> 3;
                lor(&fl[w], (1u << (c[k] & 3u)) << ((st & 7u) * 4u));
                lor(&bm[w >> 5], 1u << (w & 31u));
            }
            j0 += (uint64_t)FS_MK * TPB;
            if (j0 >= m1) break;
            load(j0, m1);
        }
        __syncthreads();
        load(n0, n1);  // in flight during the scan
        const uint64_t gw0 = (uint64_t)r * rw + t32;  // this thread's first flag word in the segment
        const uint32_t mine = threadIdx.x < nbm ? bm[threadIdx.x] : 0u;
        uint32_t out = 0, all = 0;
        for (uint32_t m = mine; m;) {
            const int w = __ffs(m) - 1;
            m &= m - 1u;
            const uint32_t f = fl[t32 + (uint32_t)w];
            const uint32_t live = fs_live(f, a.odd, gw0 + (uint32_t)w);
            all += (uint32_t)__popc(live);
            out += (uint32_t)__popc(live & ~nib_full(f, a.ops));
        }
        if (!WRITE) {
            __shared__ uint32_t s_all[NW];
            const uint32_t wo = wave_sum(out), wa = wave_sum(all);
            if (lane == 0) {
                s_w[wave] = wo;
                s_all[wave] = wa;
            }
            __syncthreads();
            if (threadIdx.x < 64) {
                const uint32_t so = wave_sum(lane < NW ? s_w[lane] : 0u);
                const uint32_t sa = wave_sum(lane < NW ? s_all[lane] : 0u);
                if (threadIdx.x == 0) {
                    a.cnt_out[r] = so;
                    a.cnt_all[r] = sa;
                }
            }
        } else {
            // workgroup exclusive scan of the per-thread entry counts
            const uint32_t incl = wave_incl_scan(out, lane);
            if (lane == 63) s_w[wave] = incl;
            __syncthreads();
            const uint32_t wt = lane < NW ? s_w[lane] : 0u;
            const uint32_t wi = wave_incl_scan(wt, lane);
            const uint32_t before = __shfl(wi - wt, wave, 64);
            uint64_t pos = off + before + incl - out;
            // one entry per iteration; a word is fetched when the last one is used up
            uint32_t m = mine, kb = 0, f = 0, wbase = 0;
            for (;;) {
                if (!kb) {
                    if (!m) break;
                    const int w = __ffs(m) - 1;
                    m &= m - 1u;
                    f = fl[t32 + (uint32_t)w];
                    kb = keep_of(f, gw0 + (uint32_t)w);
                    wbase = (uint32_t)((gw0 + (uint32_t)w) * 8u);
                    if (!kb) continue;
                }
                const int b = __ffs(kb) - 1;  // bit 4i of nibble i
                kb &= kb - 1u;
                a.out_idx[pos] = wbase + (uint32_t)(b >> 2);
                a.out_ops[pos] = (uint8_t)((f >> b) & 15u);
                ++pos;
            }
        }
        // this thread's words are read only by this thread: clear them for the next region
        for (uint32_t m = mine; m;) {
            const int w = __ffs(m) - 1;
            m &= m - 1u;
            fl[t32 + (uint32_t)w] = 0u;
        }
        if (threadIdx.x < nbm) bm[threadIdx.x] = 0u;
        m0 = n0;
        m1 = n1;
        off = noff;
    }
}

// Sums of the per-region counts: counters[1] = entries, [5] = flag words of
// the regions with marks.  One block.
template <int TPB>
__global__ __launch_bounds__(TPB) void k_fs_region_totals(const uint32_t* cnt_all, const unsigned long long* rstart,
                                                          uint32_t nregions, uint32_t rw,
                                                          unsigned long long* counters) {
    unsigned long long all = 0, scanned = 0;
    for (uint32_t r = threadIdx.x; r < nregions; r += TPB) {
        all += cnt_all[r];
        scanned += rstart[r + 1] > rstart[r] ? rw : 0u;
    }
    block_add3<TPB>(all, scanned, 0ull, &counters[1], &counters[5], nullptr);
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
