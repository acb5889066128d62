// Three-bit BFS layer kernels of 10..12 pancakes with cross-tile child lists
// (PS_CROSS_LISTS); the list geometry is in pancake_lists.h.
//
//   k_pkl_expand  scan the frontier of each 9!-rank tile; flips of <= 9
//                 elements are OR-ed into the tile's LDS bitmap (stored into
//                 next once per tile), flips of >= 10 elements are appended
//                 to the source tile's list for the child's tile (fill count
//                 in LDS: no global atomics, entries of a list are written
//                 in order so L2 merges them into whole lines)
//   k_pkl_apply   for each destination tile, gather every list that reaches
//                 it into an LDS bitmap and OR that into next
// The finalize and bookkeeping kernels of the three-bit search follow
// unchanged (there is no byte-mark array to fold).
//
// Compared with one scattered byte store per cross child (a random HBM
// write), a list entry costs 4 bytes written in whole lines and 4 bytes read
// back in order.
#pragma once

#include "bfs_kernels.h"
#include "pancake_lists.h"

namespace psk {

// LDS of k_pkl_expand: tile bitmap, child tables (pancake_lists.h), compaction
// queues, list fill counts, touched bits
constexpr size_t PKL_TAB_BYTES = 720 * sizeof(pkl::Tab6) + 720 * 4 + 504 * 4;
// Per-wave parent queue.  A full queue is expanded with lane j taking entries
// j*S .. j*S+S-1 (S = PKL_QCAP / 64): the queue holds consecutive frontier
// states, and short flips keep a child next to its parent, so lanes S
// parents apart OR into different LDS words instead of serializing on one.
constexpr int PKL_QCAP = 384;

template <int N>
inline size_t pkl_expand_lds(int tpb, uint32_t nseg) {
    size_t b = (size_t)2 * pkl::kTileWords * 4 + PKL_TAB_BYTES + (size_t)(tpb / 64) * PKL_QCAP * 4 +
               (size_t)pkl::Geo<N>::NSLOT * 4;
    if (nseg > 1 && nseg <= TOUCH_LDS_MAX) b += (size_t)((nseg + 31) / 32) * 4;
    return b;
}

// Work items of a gather: ITEM list entries of one list, read as 16-byte
// vectors (list starts are 16-byte aligned and every capacity is a multiple
// of 16, so a vector that starts below the fill count stays inside its list);
// a wave reads four vectors per lane at a time: one item of 1024 entries, or
// (1024-thread blocks, whose LDS holds the longer item table) four items of
// 256 entries, mostly of four different lists -- a flip-12 list holds at most
// 6! entries, and more lists in flight hide more of the memory latency.  A
// tile has at most 3*9!/ITEM + NSLOT work items.
template <int TPB>
constexpr uint32_t pkl_item() {
    return TPB >= 1024 ? 256u : 1024u;
}
template <int N, uint32_t ITEM>
constexpr uint32_t pkl_max_items() {
    return (uint32_t)(pkl::Geo<N>::CROSS * pkl::kTile / ITEM) + pkl::Geo<N>::NSLOT + 1;
}

// Bytes of the per-list metadata of a gather: start, count, first work item,
// and the list of every work item.
template <int N, uint32_t ITEM>
constexpr size_t pkl_gather_meta_bytes() {
    constexpr uint32_t NS = pkl::Geo<N>::NSLOT;
    return (size_t)NS * 8 + (size_t)NS * 4 + (size_t)(NS + 1) * 4 + (size_t)pkl_max_items<N, ITEM>() * 2;
}
template <int N>
inline size_t pkl_gather_meta_bytes(int tpb) {
    return tpb >= 1024 ? pkl_gather_meta_bytes<N, pkl_item<1024>()>() : pkl_gather_meta_bytes<N, pkl_item<512>()>();
}

// LDS of k_pkl_apply: tile bitmap and the gather metadata
template <int N>
inline size_t pkl_apply_lds(int tpb) {
    return (size_t)2 * pkl::kTileWords * 4 + pkl_gather_meta_bytes<N>(tpb);
}

// LDS of k_pkl_step: tile bitmap, (1024-thread blocks) the tile's new
// frontier words, child tables, the parent queues (expansion) sharing their
// space with the gather metadata (gather), list fill counts.
template <int N>
inline size_t pkl_step_lds(int tpb) {
    const size_t queues = (size_t)(tpb / 64) * PKL_QCAP * 4;
    const size_t meta = pkl_gather_meta_bytes<N>(tpb);
    const size_t front = tpb >= 1024 ? (size_t)pkl::kTileWords * 8 : 0;
    return (size_t)2 * pkl::kTileWords * 4 + front + PKL_TAB_BYTES + (queues > meta ? queues : meta) +
           (size_t)pkl::Geo<N>::NSLOT * 4;
}

// LDS pointers of the tile expansion.
struct PklExpLds {
    uint32_t* win;  // tile bitmap (next)
    const pkl::Tab6* tab6;
    const uint32_t* tab6lo;
    const uint32_t* tab3;
    uint32_t* q;    // this wave's parent queue
    uint32_t* cnt;  // list fill counts of the tile
};

// Expands the parents of tile T whose bits are in frontier words [pw0, pw1)
// (tile words start at w0; front(i) reads frontier word i): flips of <= 9
// elements into lds.win, flips of >= 10 elements appended to the tile's lists.
// Wave-synchronous only; the caller brackets it with block barriers.
template <int N, int TPB, class Front>
__device__ __forceinline__ void pkl_expand_range(Front front, uint64_t w0, uint64_t pw0,
                                                 uint64_t pw1, const PklExpLds& lds, const pkl::TileK& tk,
                                                 uint32_t* lists_t, unsigned long long& parents, uint32_t& err,
                                                 bool& any) {
    constexpr int NW = TPB / 64;
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    uint32_t* win = lds.win;
    uint32_t* q = lds.q;
    uint32_t* cnt = lds.cnt;
    constexpr int CROSS = pkl::Geo<N>::CROSS;
    auto expand_one = [&](uint32_t off, const pkl::TabRow& row) {
        ++parents;
        any = true;
        uint32_t xs[CROSS], xc[CROSS];
        pkl::children_row<N>(
            row, tk, off, [&](uint32_t c) { lor(&win[c >> 5], 1u << (c & 31)); },
            [&](int k, uint32_t slot, uint32_t child) {
                xs[k - 10] = slot;
                xc[k - 10] = child;
            });
        // the fill-count atomics of the cross flips back to back, under one wait
        // (a slot out of range, which slot_of never returns, counts on the last
        // slot and is reported, its child not stored)
        uint32_t pos[CROSS];
#pragma unroll
        for (int f = 0; f < CROSS; ++f) {
            const uint32_t sl = xs[f] < pkl::kSlots[f] ? xs[f] : pkl::kSlots[f] - 1u;
            pos[f] = __hip_atomic_fetch_add(&cnt[pkl::kSlotBase[f] + sl], 1u, __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_WORKGROUP);
        }
#pragma unroll
        for (int f = 0; f < CROSS; ++f) {
            if (xs[f] < pkl::kSlots[f] && pos[f] < pkl::kCap[f])
                lists_t[(uint32_t)f * pkl::kTile + xs[f] * pkl::kCap[f] + pos[f]] = xc[f];
            else
                err |= 2u;
        }
    };
    auto row_of = [&](uint32_t off) { return pkl::tab_row(lds.tab6, lds.tab6lo, lds.tab3, off); };
    auto run_batch = [&](uint32_t n) {
        if (n == (uint32_t)PKL_QCAP) {
            constexpr uint32_t S = PKL_QCAP / 64;
            uint32_t off = q[lane * S];
            pkl::TabRow row = row_of(off);
#pragma unroll 1
            for (uint32_t r = 0; r < S; ++r) {
                // the next parent's table row is read before this parent is
                // expanded (the last step reads its own row again)
                const uint32_t off_n = q[lane * S + (r + 1 < S ? r + 1 : r)];
                const pkl::TabRow row_n = row_of(off_n);
                expand_one(off, row);
                off = off_n;
                row = row_n;
            }
        } else {
            for (uint32_t j = lane; j < n; j += 64) expand_one(q[j], row_of(q[j]));
        }
    };
    int qn = 0;  // wave-uniform queue fill
    for (uint64_t g = pw0 + (uint64_t)wave * 64; g < pw1; g += (uint64_t)NW * 64) {
        const uint64_t wi = g + lane;
        uint64_t bits = (wi < pw1) ? front(wi) : 0ull;
        const uint32_t c = (uint32_t)__popcll(bits);
        const uint32_t incl = wave_incl_scan(c, lane);
        const int total = (int)__shfl(incl, 63, 64);
        int off = qn + (int)(incl - c);
        int end_all = qn + total;
        const uint32_t base = (uint32_t)((wi - w0) * 64);
        for (;;) {
            while (bits && off < PKL_QCAP) {
                q[off++] = base + (uint32_t)(__ffsll((unsigned long long)bits) - 1);
                bits &= bits - 1;
            }
            if (end_all < PKL_QCAP) {
                qn = end_all;
                break;
            }
            wave_sync();
            run_batch(PKL_QCAP);
            wave_sync();
            off -= PKL_QCAP;
            end_all -= PKL_QCAP;
            qn = 0;
            if (end_all == 0) break;
        }
    }
    wave_sync();
    run_batch((uint32_t)qn);
}

// Stores the tile's list fill counts and flags every destination tile of a
// non-empty list.  all_counts: also store zero counts.
template <int N, int TPB>
__device__ __forceinline__ void pkl_store_counts(uint32_t T, const uint32_t* cnt, uint32_t* counts, uint8_t* flags,
                                                 bool all_counts) {
    using G = pkl::Geo<N>;
    for (uint32_t i = threadIdx.x; i < G::NSLOT; i += TPB) {
        const uint32_t c = cnt[i];
        if (c || all_counts) counts[(uint64_t)T * G::NSLOT + i] = c;
        if (c) {
            const int k = i < pkl::kSlotBase[1] ? 10 : (i < pkl::kSlotBase[2] ? 11 : 12);
            flags[pkl::dest_of<N>(T, k, i - pkl::kSlotBase[k - 10])] = 1;
        }
    }
}

// Gathers every list that reaches destination tile t into win (which the
// caller has cleared), clearing the lists' fill counts.  Every thread of the
// block calls it; it ends with a block barrier.  meta: LDS of
// pkl_gather_meta_bytes<N, pkl_item<TPB>()>() bytes, 8-byte aligned.
template <int N, int TPB>
__device__ __forceinline__ void pkl_gather(uint32_t t, uint32_t* counts, const uint32_t* lists, uint32_t* win,
                                           uint8_t* meta, uint32_t& err) {
    using G = pkl::Geo<N>;
    constexpr int NW = TPB / 64;
    constexpr uint32_t ITEM = pkl_item<TPB>();
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    unsigned long long* lbase = reinterpret_cast<unsigned long long*>(meta);
    uint32_t* lcnt = reinterpret_cast<uint32_t*>(lbase + G::NSLOT);
    uint32_t* lstart = lcnt + G::NSLOT;  // NSLOT + 1 entries
    uint16_t* item_list = reinterpret_cast<uint16_t*>(lstart + G::NSLOT + 1);
    for (uint32_t i = tid; i < G::NSLOT; i += TPB) {
        int k = 10;
        uint32_t src = 0, slot = 0;
        pkl::incoming<N>(t, i, &k, &src, &slot);
        const uint64_t ci = (uint64_t)src * G::NSLOT + pkl::kSlotBase[k - 10] + slot;
        uint32_t c = counts[ci];
        if (c) {
            counts[ci] = 0;
            if (c > pkl::kCap[k - 10]) {
                err |= 2u;
                c = 0;
            }
        }
        lcnt[i] = c;
        lbase[i] = pkl::list_base(src, G::CROSS, k, slot);
    }
    __syncthreads();
    // work items: exclusive scan over the lists (wave 0)
    if (wave == 0) {
        uint32_t run = 0;
        for (uint32_t b = 0; b < G::NSLOT; b += 64) {
            const uint32_t i = b + lane;
            const uint32_t c = i < G::NSLOT ? (lcnt[i] + ITEM - 1) / ITEM : 0u;
            const uint32_t incl = wave_incl_scan(c, lane);
            if (i < G::NSLOT) lstart[i] = run + incl - c;
            run += __shfl(incl, 63, 64);
        }
        if (lane == 0) lstart[G::NSLOT] = run;
    }
    __syncthreads();
    const uint32_t total = lstart[G::NSLOT];
    // the list of every work item
    for (uint32_t i = tid; i < G::NSLOT; i += TPB)
        for (uint32_t w = lstart[i]; w < lstart[i + 1]; ++w) item_list[w] = (uint16_t)i;
    __syncthreads();
    const uint32_t tbase = t * pkl::kTile;
    constexpr int UPI = (int)(ITEM / 256);  // vectors per lane per item
    constexpr int GRP = 4 / UPI;            // items per wave and step
    // applies vectors v[0..UPI) of the item that starts at entry e0 of a list of n
    auto apply = [&](const uint4* v, uint32_t e0, uint32_t n) {
        if (e0 + ITEM <= n) {
            // a full item, branch-free: an entry outside the tile ORs no bit
#pragma unroll
            for (int u = 0; u < UPI; ++u) {
                const uint32_t ev[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const uint32_t o = ev[m] - tbase;
                    const bool ok = o < pkl::kTile;
                    err |= ok ? 0u : 4u;
                    lor(&win[ok ? (o >> 5) : (uint32_t)lane], ok ? (1u << (o & 31)) : 0u);
                }
            }
        } else {
            // the last item of a list: only the groups of 256 entries that hold
            // any, each entry tested without a branch (an entry past the end
            // ORs no bit)
#pragma unroll
            for (int u = 0; u < UPI; ++u) {
                if (e0 + 256u * u >= n) break;  // wave-uniform
                const uint32_t e = e0 + 4u * (lane + 64u * u);
                const uint32_t ev[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const uint32_t o = ev[m] - tbase;
                    const bool live = e + m < n;
                    const bool ok = live && o < pkl::kTile;
                    err |= (live && !ok) ? 4u : 0u;
                    lor(&win[ok ? (o >> 5) : (uint32_t)lane], ok ? (1u << (o & 31)) : 0u);
                }
            }
        }
    };
    // wave w takes items w, w + NW, .. w + (GRP-1)*NW, then moves on by GRP*NW;
    // every load of a step is issued before its first item is applied
    for (uint32_t w = wave; w < total; w += (uint32_t)(GRP * NW)) {
        uint4 v[4];
        uint32_t e0[GRP], n[GRP];
#pragma unroll
        for (int g = 0; g < GRP; ++g) {
            const uint32_t wi = w + (uint32_t)(g * NW);
            e0[g] = 0;
            n[g] = 0;
            if (wi < total) {  // wave-uniform
                // the item's list, start and length are wave-uniform: scalar registers
                const uint32_t li = __builtin_amdgcn_readfirstlane(item_list[wi]);
                e0[g] = __builtin_amdgcn_readfirstlane((wi - lstart[li]) * ITEM);
                n[g] = __builtin_amdgcn_readfirstlane(lcnt[li]);
                const unsigned long long lb = lbase[li];
                const uint64_t first = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(lb >> 32)) << 32) |
                                       __builtin_amdgcn_readfirstlane((uint32_t)lb);
                const uint4* src = reinterpret_cast<const uint4*>(lists + first + e0[g]);
#pragma unroll
                for (int u = 0; u < UPI; ++u) {
                    const uint32_t e = e0[g] + 4u * (lane + 64u * u);
                    v[g * UPI + u] = (e < n[g]) ? src[lane + 64 * u] : make_uint4(0u, 0u, 0u, 0u);
                }
            } else {
#pragma unroll
                for (int u = 0; u < UPI; ++u) v[g * UPI + u] = make_uint4(0u, 0u, 0u, 0u);
            }
        }
#pragma unroll
        for (int g = 0; g < GRP; ++g)
            if (n[g]) apply(v + g * UPI, e0[g], n[g]);
    }
    __syncthreads();
}

// Takes the next of n tiles (~0 when none is left): a plain read first, so
// blocks arriving after the last tile make no atomic.  Thread 0 only.
// After a split finalize: the layer just finalized (its count in acc_count)
// is expanded bottom-up when it outnumbers the states not yet visited, which
// bound the bottom-up work (Beamer's direction-optimizing BFS).  Every block
// of the launches after the finalize reads the same values.
__device__ __forceinline__ bool pkl_pull_now(const BfsDevState* st, uint64_t nbits) {
    const unsigned long long cnt = st->acc_count, vis = st->visited + cnt;
    return st->direction == 1u && cnt > (nbits > vis ? nbits - vis : 0ull);
}

__device__ __forceinline__ unsigned long long pkl_take_tile(unsigned long long* counter, unsigned long long n) {
    const unsigned long long taken = __hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (taken >= n) return ~0ull;
    const unsigned long long k = atomicAdd(counter, 1ull);  // other blocks may have taken the rest
    return k < n ? k : ~0ull;
}

template <int N, int TPB>
__global__ __launch_bounds__(TPB) void k_pkl_expand(ExpandArgs a, ListArgs x) {
    using G = pkl::Geo<N>;
    constexpr int NW = TPB / 64;
    constexpr uint32_t WIN = 2 * pkl::kTileWords;  // u32 words of a tile bitmap
    extern __shared__ uint32_t smem[];
    BfsDevState* st = a.st;
    if (st->done) return;
    const int tid = threadIdx.x, wave = tid / 64;
    PklExpLds lds;
    lds.win = smem;
    pkl::Tab6* tab6 = reinterpret_cast<pkl::Tab6*>(smem + WIN);  // WIN * 4 bytes is a multiple of 16
    uint32_t* tab6lo = reinterpret_cast<uint32_t*>(tab6 + 720);
    uint32_t* tab3 = tab6lo + 720;
    lds.tab6 = tab6;
    lds.tab6lo = tab6lo;
    lds.tab3 = tab3;
    lds.q = tab3 + 504 + wave * PKL_QCAP;
    lds.cnt = tab3 + 504 + NW * PKL_QCAP;
    uint32_t* tl = lds.cnt + G::NSLOT;
    const uint32_t tl_words = (a.nseg > 1 && a.nseg <= TOUCH_LDS_MAX) ? (a.nseg + 31) / 32 : 0u;
    const bool track = a.nseg > 1;
    const uint32_t L = st->layer;
    const uint64_t* __restrict__ cur = a.bits[L % 3];
    uint32_t* nxt = reinterpret_cast<uint32_t*>(a.bits[(L + 1) % 3]);
    const unsigned long long* sc_cur = a.seg_count + (uint64_t)(L % 3) * a.nseg;
    __shared__ unsigned long long s_tile;
    __shared__ uint32_t s_local;
    for (uint32_t i = tid; i < tl_words; i += TPB) tl[i] = 0;
    bool tables = false;  // built before the block's first live tile

    unsigned long long parents = 0, scan = 0;
    uint32_t err = 0;
    uint64_t iter = 0;
    for (;;) {
        if (tid == 0) {
            if (a.persistent) {
                const unsigned long long k = pkl_take_tile(&st->tile_counter, a.tile_end - a.tile_begin);
                s_tile = k == ~0ull ? ~0ull : a.tile_begin + k;
            } else {
                s_tile = (iter == 0) ? a.tile_begin + blockIdx.x : ~0ull;
            }
        }
        __syncthreads();
        const uint64_t tile = s_tile;
        ++iter;
        if (tile >= a.tile_end) break;
        const uint64_t w0 = tile * pkl::kTileWords, w1 = w0 + pkl::kTileWords;
        const uint64_t pw0 = w0 > a.word_lo ? w0 : a.word_lo;
        const uint64_t pw1 = w1 < a.word_hi ? w1 : a.word_hi;
        bool live = pw0 < pw1;
        if (live && track) {
            const uint64_t s0 = (pw0 * 64) >> a.seg_shift, s1 = (pw1 * 64 - 1) >> a.seg_shift;
            live = false;
            for (uint64_t s = s0; s <= s1; ++s) live |= sc_cur[s] != 0;
        }
        if (!live) {
            __syncthreads();
            continue;
        }
        const uint32_t T = (uint32_t)tile;
        if (tid == 0) s_local = 0;
        if (!tables) {
            for (uint32_t l = tid; l < 720; l += TPB) tab6[l] = pkl::tab6_entry(l, &tab6lo[l]);
            for (uint32_t h = tid; h < 504; h += TPB) tab3[h] = pkl::tab3_entry(h);
            tables = true;
        }
        for (uint32_t i = tid; i < WIN; i += TPB) lds.win[i] = 0;
        // fill counts start where an earlier piece of this tile (plain mode)
        // left them; the apply pass of the previous layer cleared the rest
        for (uint32_t i = tid; i < G::NSLOT; i += TPB) lds.cnt[i] = x.counts[(uint64_t)T * G::NSLOT + i];
        uint32_t dp[3];
        pkl::tile_digits<N>(T, dp);
        const pkl::TileK tk = pkl::tile_consts<N>(dp);
        uint32_t* lists_t = x.lists + (uint64_t)T * G::CROSS * pkl::kTile;
        __syncthreads();
        if (tid == 0) scan += pw1 - pw0;
        bool any = false;
        pkl_expand_range<N, TPB>([&](uint64_t i) { return cur[i]; }, w0, pw0, pw1, lds, tk, lists_t, parents, err,
                                 any);
        if (any) s_local = 1;
        __syncthreads();
        // a tile expanded whole in one visit (batched mode) owns its range of
        // next until the apply pass; a piece of a tile (plain mode) ORs
        uint32_t* dst = nxt + w0 * 2;
        const bool own = pw0 == w0 && pw1 == w1;
        for (uint32_t i = tid; i < WIN; i += TPB) {
            const uint32_t v = lds.win[i];
            if (v) {
                if (own) dst[i] = v;
                else gor(&dst[i], v);
            }
        }
        pkl_store_counts<N, TPB>(T, lds.cnt, x.counts, x.flags, true);
        if (tid == 0 && track && s_local) {
            const uint64_t s0 = (w0 * 64) >> a.seg_shift, s1 = (w1 * 64 - 1) >> a.seg_shift;
            for (uint64_t s = s0; s <= s1; ++s) {
                if (tl_words) lor(&tl[s >> 5], 1u << (s & 31));
                else gor(&a.touched[s >> 5], 1u << (s & 31));
            }
        }
        __syncthreads();
    }
    __syncthreads();
    for (uint32_t i = tid; i < tl_words; i += TPB) {
        const uint32_t v = tl[i];
        if (v) gor(&a.touched[i], v);
    }
    if (err) atomicOr(&st->error, err);
    block_add3<TPB>(parents * (unsigned long long)(N - 1), parents, scan, &st->children, &st->expanded,
                    &st->scan_words);
}

template <int N, int TPB>
__global__ __launch_bounds__(TPB) void k_pkl_apply(ApplyArgs a) {
    constexpr uint32_t WIN = 2 * pkl::kTileWords;
    extern __shared__ uint32_t smem[];
    BfsDevState* st = a.st;
    if (st->done) return;
    const int tid = threadIdx.x;
    uint32_t* win = smem;
    uint8_t* meta = reinterpret_cast<uint8_t*>(smem + WIN);  // WIN is even: 8-byte aligned
    const uint32_t L = st->layer;
    uint32_t* nxt = reinterpret_cast<uint32_t*>(a.bits[(L + 1) % 3]);
    __shared__ unsigned long long s_tile;
    __shared__ uint32_t s_flag;
    uint32_t err = 0;
    uint64_t iter = 0;
    for (;;) {
        if (tid == 0) {
            // destination tiles carry similar work: a static stride, no shared counter
            if (a.persistent) s_tile = a.tile_begin + blockIdx.x + iter * gridDim.x;
            else s_tile = (iter == 0) ? a.tile_begin + blockIdx.x : ~0ull;
            s_flag = s_tile < a.tile_end ? a.x.flags[s_tile] : 0u;
        }
        __syncthreads();
        const uint64_t tile = s_tile;
        ++iter;
        if (tile >= a.tile_end) break;
        if (!s_flag) {
            __syncthreads();
            continue;
        }
        const uint32_t t = (uint32_t)tile;
        if (tid == 0) a.x.flags[t] = 0;
        for (uint32_t i = tid; i < WIN; i += TPB) win[i] = 0;
        pkl_gather<N, TPB>(t, a.x.counts, a.x.lists, win, meta, err);
        uint32_t* dst = nxt + (uint64_t)t * WIN;
        for (uint32_t i = tid; i < WIN; i += TPB) {
            const uint32_t vv = win[i];
            if (vv) dst[i] |= vv;
        }
        if (tid == 0 && a.nseg > 1) {
            const uint64_t s0 = ((uint64_t)t * pkl::kTile) >> a.seg_shift;
            const uint64_t s1 = ((uint64_t)(t + 1) * pkl::kTile - 1) >> a.seg_shift;
            for (uint64_t sg = s0; sg <= s1; ++sg) gor(&a.touched[sg >> 5], 1u << (sg & 31));
        }
        __syncthreads();
    }
    if (err) atomicOr(&st->error, err);
}

// One whole layer of the batched search with cross-tile lists, per tile:
//  1. (not in the first launch) finalize the next layer of the tile: gather
//     the lists that reach it (produced by the previous launch), then
//     new = (next | gathered) & ~old & ~cur, count, hash, old = 0 -- or only
//     clear old where nothing reached the tile;
//  2. expand the tile's new frontier: flips of <= 9 elements into the tile's
//     bitmap, stored into the array cleared in step 1 (the following layer's
//     next), flips of >= 10 elements into the lists of the other parity.
// Every dependency between the two steps is inside the tile; across tiles they
// meet only through lists completed by the previous launch.  The launch visits
// the st->active_tiles tiles of a.active (k_pkl_layer_end lists the tiles
// that have anything to finalize or clear); the first launch (first = 1)
// expands the start layer, in a.first_tile.  Layer L (st->layer) is the last
// finalized layer; launch m finalizes and expands layer m = L + 1.
// Four waves per SIMD (16 per CU): one 1024-thread block or two 512-thread
// blocks per CU, as the LDS allows (at most 128 VGPRs; unbounded, the compiler
// takes more and leaves one 512-thread block per CU).  1024-thread blocks keep
// the finalized frontier words in LDS and expand from there.
template <int N, int TPB, int ROLE = 0>
__global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(4))) void k_pkl_step(StepArgs a) {
    using G = pkl::Geo<N>;
    constexpr int NW = TPB / 64;
    constexpr uint32_t WIN = 2 * pkl::kTileWords;
    constexpr bool FRONT_LDS = TPB >= 1024;
    static_assert(pkl::kTileWords % 2 == 0, "tile words are finalized in pairs");
    extern __shared__ uint32_t smem[];
    BfsDevState* st = a.st;
    if (st->done) return;
    const bool first = a.first != 0;
    // split layers: the role-0 launch only finalizes; a role-1 launch then
    // expands top-down unless the pull launch expands bottom-up
    const bool split = !first && st->split != 0u;
    constexpr bool expand_only = ROLE == 1;
    if (expand_only && (!split || pkl_pull_now(st, a.nbits))) return;
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    PklExpLds lds;
    lds.win = smem;
    // WIN * 4 bytes is a multiple of 16: every region below is 16-byte aligned
    uint64_t* flds = FRONT_LDS ? reinterpret_cast<uint64_t*>(smem + WIN) : nullptr;
    pkl::Tab6* tab6 = reinterpret_cast<pkl::Tab6*>(smem + WIN + (FRONT_LDS ? WIN : 0));
    uint32_t* tab6lo = reinterpret_cast<uint32_t*>(tab6 + 720);
    uint32_t* tab3 = tab6lo + 720;
    lds.tab6 = tab6;
    lds.tab6lo = tab6lo;
    lds.tab3 = tab3;
    uint32_t* shared_region = tab3 + 504;  // queues (expansion) or gather metadata
    lds.q = shared_region + wave * PKL_QCAP;
    {
        const size_t qbytes = (size_t)NW * PKL_QCAP * 4, mbytes = pkl_gather_meta_bytes<N, pkl_item<TPB>()>();
        lds.cnt = shared_region + (qbytes > mbytes ? qbytes : mbytes) / 4;
    }
    uint8_t* meta = reinterpret_cast<uint8_t*>(shared_region);  // tab3 + 504 words: 8-byte aligned
    const uint32_t L = st->layer;
    const uint32_t nvisit = first ? (a.first_tile < a.tiles ? 1u : 0u) : st->active_tiles;
    const uint64_t* cur = a.bits[L % 3];
    uint64_t* nw = a.bits[(L + 1) % 3];
    uint64_t* old = a.bits[(L + 2) % 3];
    // the expansion: frontier, its next array, list parity
    const uint64_t* front = first ? a.bits[0] : nw;
    const bool front_lds = FRONT_LDS && !first && !expand_only;
    uint32_t* enext = reinterpret_cast<uint32_t*>(first ? a.bits[1] : old);
    const int gp = (int)(L & 1u), xp = first ? 0 : (int)((L + 1u) & 1u);
    uint32_t* tc_new = a.tcount + (uint64_t)(first ? 0 : (L + 1) % 3) * a.tiles;
    const uint32_t* tc_old = a.tcount + (uint64_t)((L + 2) % 3) * a.tiles;
    __shared__ unsigned long long s_tile;
    __shared__ uint32_t s_xf, s_wf, s_oldc, s_local;
    __shared__ unsigned long long s_red[2][NW];
    __shared__ unsigned long long s_cnt;
    // per-block tallies kept by thread 0 (scan, fin_full, fin_clear, count, hash):
    // in LDS rather than in registers every thread would carry across the tile loop
    __shared__ unsigned long long s_tally[5];
    if (tid < 5) s_tally[tid] = 0;
    bool tables = false;
    unsigned long long parents = 0;
    uint32_t err = 0;
    // the first launch has no bookkeeping launch after it to reset tile_counter;
    // an expand-only launch follows the finalize launch of the same layer
    unsigned long long* counter = (first || expand_only) ? &st->chunk_counter : &st->tile_counter;
    for (;;) {
        if (tid == 0) {
            const unsigned long long k = pkl_take_tile(counter, nvisit);
            s_tile = k == ~0ull ? ~0ull : (first ? (unsigned long long)a.first_tile : (unsigned long long)a.active[k]);
            if (s_tile != ~0ull) {
                s_xf = first ? 0u : a.xflags[gp][s_tile];
                s_wf = first ? 0u : a.wflags[gp][s_tile];
                s_oldc = first ? 0u : tc_old[s_tile];
            }
        }
        __syncthreads();
        const uint64_t tile = s_tile;
        if (tile >= a.tiles) break;
        const uint32_t T = (uint32_t)tile;
        const uint64_t w0 = tile * pkl::kTileWords, w1 = w0 + pkl::kTileWords;
        bool expand = true;
        if (expand_only) {
            // the tile's frontier count, written by the finalize launch
            expand = tc_new[T] != 0u;
        } else if (!first) {
            // 1. finalize the tile's next layer
            unsigned long long c = 0, h = 0;
            if (s_xf | s_wf) {
                for (uint32_t i = tid; i < WIN; i += TPB) lds.win[i] = 0;
                __syncthreads();
                if (s_xf) pkl_gather<N, TPB>(T, a.counts[gp], a.lists[gp], lds.win, meta, err);
                // pairs of words (w0 is even: 16-byte aligned), one pair per
                // thread and step; more in flight push the kernel past 128 VGPRs
                const ulonglong2* gw = reinterpret_cast<const ulonglong2*>(lds.win);
                ulonglong2* nw2 = reinterpret_cast<ulonglong2*>(nw + w0);
                ulonglong2* old2 = reinterpret_cast<ulonglong2*>(old + w0);
                const ulonglong2* cur2 = reinterpret_cast<const ulonglong2*>(cur + w0);
                ulonglong2* vis2 = a.visited ? reinterpret_cast<ulonglong2*>(a.visited + w0) : nullptr;
                for (uint32_t q = tid; q < pkl::kTileWords / 2; q += TPB) {
                    const ulonglong2 vn = nw2[q], vo = old2[q];
                    const ulonglong2 vc = a.odd ? cur2[q] : make_ulonglong2(0ull, 0ull);
                    const ulonglong2 g = gw[q];
                    ulonglong2 v;
                    v.x = (vn.x | g.x) & ~vo.x & ~vc.x;
                    v.y = (vn.y | g.y) & ~vo.y & ~vc.y;
                    nw2[q] = v;
                    old2[q] = make_ulonglong2(0ull, 0ull);
                    if (vis2 && (v.x | v.y)) {
                        const ulonglong2 vv = vis2[q];
                        vis2[q] = make_ulonglong2(vv.x | v.x, vv.y | v.y);
                    }
                    if (FRONT_LDS) reinterpret_cast<ulonglong2*>(flds)[q] = v;
                    c += (unsigned long long)(__popcll(v.x) + __popcll(v.y));
                    if (a.hash) h += word_hash(w0 + 2 * q, v.x) + word_hash(w0 + 2 * q + 1, v.y);
                }
                if (tid == 0) {
                    a.xflags[gp][T] = 0;
                    a.wflags[gp][T] = 0;
                    s_tally[1] += pkl::kTileWords;
                }
            } else if (s_oldc) {
                for (uint64_t j = tid; j < pkl::kTileWords; j += TPB) old[w0 + j] = 0;
                if (tid == 0) s_tally[2] += pkl::kTileWords;
            }
            c = wave_sum(c);
            h = wave_sum(h);
            if (lane == 0) {
                s_red[0][wave] = c;
                s_red[1][wave] = h;
            }
            __syncthreads();
            if (tid == 0) {
                unsigned long long sc = 0, sh = 0;
                for (int w = 0; w < NW; ++w) {
                    sc += s_red[0][w];
                    sh += s_red[1][w];
                }
                tc_new[T] = (uint32_t)sc;
                s_tally[3] += sc;
                s_tally[4] += sh;
                s_cnt = sc;
            }
            __syncthreads();
            expand = s_cnt != 0 && !split;
        }
        if (expand) {
            // 2. expand the tile's frontier
            if (tid == 0) s_local = 0;
            if (!tables) {
                for (uint32_t l = tid; l < 720; l += TPB) tab6[l] = pkl::tab6_entry(l, &tab6lo[l]);
                for (uint32_t hh = tid; hh < 504; hh += TPB) tab3[hh] = pkl::tab3_entry(hh);
                tables = true;
            }
            for (uint32_t i = tid; i < WIN; i += TPB) lds.win[i] = 0;
            // the lists of this parity were consumed (counts cleared) by the previous launch
            for (uint32_t i = tid; i < G::NSLOT; i += TPB) lds.cnt[i] = 0;
            uint32_t dp[3];
            pkl::tile_digits<N>(T, dp);
            const pkl::TileK tk = pkl::tile_consts<N>(dp);
            uint32_t* lists_t = a.lists[xp] + (uint64_t)T * G::CROSS * pkl::kTile;
            __syncthreads();
            if (tid == 0) s_tally[0] += pkl::kTileWords;
            bool any = false;
            unsigned long long tile_parents = 0;
            pkl_expand_range<N, TPB>(
                [&](uint64_t i) { return front_lds ? flds[i - w0] : front[i]; }, w0, w0, w1, lds, tk, lists_t,
                tile_parents, err, any);
            parents += tile_parents;
            if (any) s_local = 1;
            __syncthreads();
            uint32_t* dst = enext + w0 * 2;
            for (uint32_t i = tid; i < WIN; i += TPB) {
                const uint32_t v = lds.win[i];
                if (v) dst[i] = v;
            }
            pkl_store_counts<N, TPB>(T, lds.cnt, a.counts[xp], a.xflags[xp], false);
            if (tid == 0 && s_local) a.wflags[xp][T] = 1;
            if (first) {
                // the start layer's count per tile (the finalize two launches on clears old by it)
                tile_parents = wave_sum(tile_parents);
                if (lane == 0) s_red[0][wave] = tile_parents;
                __syncthreads();
                if (tid == 0) {
                    unsigned long long sc = 0;
                    for (int w = 0; w < NW; ++w) sc += s_red[0][w];
                    a.tcount[T] = (uint32_t)sc;
                }
            }
        }
        __syncthreads();
    }
    if (err) atomicOr(&st->error, err);
    if (tid == 0) {
        if (s_tally[1]) atomicAdd(&st->fin_full_words, s_tally[1]);
        if (s_tally[2]) atomicAdd(&st->fin_clear_words, s_tally[2]);
        if (s_tally[3]) atomicAdd(&st->acc_count, s_tally[3]);
        if (s_tally[4]) atomicAdd(&st->acc_hash, s_tally[4]);
    }
    block_add3<TPB>(parents * (unsigned long long)(N - 1), parents, tid == 0 ? s_tally[0] : 0ull, &st->children,
                    &st->expanded, &st->scan_words);
}

// Bottom-up expansion of a split layer (direction-optimizing BFS): after the
// finalize launch, every state not yet visited looks for a neighbour in the
// layer just finalized -- flips are involutions, so its neighbours are its
// children.  Flips of <= 9 elements stay in the tile and are tested in an LDS
// copy of the tile's frontier words; only a state none of them reaches tests
// its flips of 10..N elements, one read of the frontier each (the list
// arithmetic gives those children as global ranks).  A state with a
// neighbour in the frontier is in the next layer (the graph has no shorter
// path to it); its bit goes to an LDS window stored into the tile's next
// words, which nothing else writes.  No lists are filled.  Each state is
// first tested against its flips of 2..6 elements, whose tile offsets the
// six-position table row gives directly; only the states none of them reaches
// (a fraction of them in the dense layers a split expands) are queued per wave
// for the full test, which computes the flips of 7..9 elements by the sweep.
constexpr int PKL_QB = 128;  // per-wave queue of states the table check leaves to the full check
template <int N>
inline size_t pkl_pull_lds(int tpb) {
    return (size_t)2 * pkl::kTileWords * 4 + (size_t)pkl::kTileWords * 8 + PKL_TAB_BYTES +
           (size_t)(tpb / 64) * (PKL_QCAP + PKL_QB) * 4;
}

template <int N, int TPB>
__global__ __launch_bounds__(TPB) void k_pkl_pull(PullArgs a) {
    using G = pkl::Geo<N>;
    constexpr int NW = TPB / 64;
    constexpr int CROSS = G::CROSS;
    constexpr uint32_t WIN = 2 * pkl::kTileWords;
    extern __shared__ uint32_t smem[];
    BfsDevState* st = a.st;
    if (st->done || st->split == 0u || !pkl_pull_now(st, a.nbits)) return;
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    uint32_t* win = smem;
    uint64_t* flds = reinterpret_cast<uint64_t*>(smem + WIN);  // WIN * 4 is a multiple of 16
    pkl::Tab6* tab6 = reinterpret_cast<pkl::Tab6*>(smem + 2 * WIN);
    uint32_t* tab6lo = reinterpret_cast<uint32_t*>(tab6 + 720);
    uint32_t* tab3 = tab6lo + 720;
    uint32_t* q = tab3 + 504 + wave * PKL_QCAP;
    uint32_t* qb = tab3 + 504 + NW * PKL_QCAP + wave * PKL_QB;
    const uint32_t L = st->layer;
    const uint64_t* front = a.bits[(L + 1) % 3];             // the layer just finalized
    uint32_t* nxt = reinterpret_cast<uint32_t*>(a.bits[(L + 2) % 3]);  // cleared by that launch
    const int xp = (int)((L + 1u) & 1u);
    __shared__ unsigned long long s_tile;
    __shared__ uint32_t s_any;
    for (uint32_t l = tid; l < 720; l += TPB) tab6[l] = pkl::tab6_entry(l, &tab6lo[l]);
    for (uint32_t hh = tid; hh < 504; hh += TPB) tab3[hh] = pkl::tab3_entry(hh);
    unsigned long long words = 0, probes = 0;
    for (;;) {
        if (tid == 0) s_tile = pkl_take_tile(&st->chunk_counter, a.tiles);
        __syncthreads();
        const uint64_t tile = s_tile;
        if (tile >= a.tiles) break;
        const uint32_t T = (uint32_t)tile;
        const uint64_t w0 = tile * pkl::kTileWords;
        for (uint32_t i = tid; i < WIN; i += TPB) win[i] = 0;
        for (uint32_t i = tid; i < pkl::kTileWords; i += TPB) flds[i] = front[w0 + i];
        if (tid == 0) s_any = 0;
        uint32_t dp[3];
        pkl::tile_digits<N>(T, dp);
        const pkl::TileK tk = pkl::tile_consts<N>(dp);
        __syncthreads();
        if (tid == 0) words += pkl::kTileWords;
        bool any = false;
        auto probe = [&](uint32_t off) {
            bool found = false;
            uint32_t xc[CROSS];
            pkl::children_row<N>(
                pkl::tab_row(tab6, tab6lo, tab3, off), tk, off,
                [&](uint32_t c) { found |= ((flds[c >> 6] >> (c & 63)) & 1ull) != 0ull; },
                [&](int k, uint32_t, uint32_t child) { xc[k - 10] = child; });
            if (!found) {
#pragma unroll
                for (int f = 0; f < CROSS; ++f) {
                    const uint64_t g = xc[f];  // the child's rank
                    if (g >= a.nbits) continue;
                    ++probes;
                    found |= ((front[g >> 6] >> (g & 63)) & 1ull) != 0ull;
                }
            }
            if (found) {
                lor(&win[off >> 5], 1u << (off & 31));
                any = true;
            }
        };
        // the table check: flips of 2..6 elements; a state it leaves goes to
        // the wave's second queue, which is drained 64 states at a time
        auto quick = [&](uint32_t off) {
            const uint32_t h = off / 720u;
            const uint32_t l = off - h * 720u;
            const pkl::Tab6 t = tab6[l];
            const uint32_t b6 = off - l;
            const uint32_t c[5] = {b6 + (t.c234 & 1023u), b6 + ((t.c234 >> 10) & 1023u), b6 + ((t.c234 >> 20) & 1023u),
                                   b6 + (t.c56 & 1023u), b6 + ((t.c56 >> 10) & 1023u)};
            bool found = false;
#pragma unroll
            for (int k = 0; k < 5; ++k) found |= ((flds[c[k] >> 6] >> (c[k] & 63u)) & 1ull) != 0ull;
            return found;
        };
        int qbn = 0;  // wave-uniform fill of qb
        auto run_queue = [&](uint32_t n) {
            for (uint32_t j0 = 0; j0 < n; j0 += 64) {
                const uint32_t j = j0 + lane;
                const bool has = j < n;
                const uint32_t off = has ? q[j] : 0u;
                const bool found = has && quick(off);
                if (found) {
                    lor(&win[off >> 5], 1u << (off & 31));
                    any = true;
                }
                const bool rest = has && !found;
                const unsigned long long m = __ballot(rest);
                if (rest) qb[qbn + __popcll(m & ((1ull << lane) - 1ull))] = off;
                qbn += __popcll(m);
                if (qbn >= 64) {
                    wave_sync();
                    probe(qb[lane]);
                    const uint32_t tail = (uint32_t)(qbn - 64) > (uint32_t)lane ? qb[64 + lane] : 0u;
                    wave_sync();
                    if ((uint32_t)(qbn - 64) > (uint32_t)lane) qb[lane] = tail;
                    qbn -= 64;
                    wave_sync();
                }
            }
        };
        // the tile's unvisited states, compacted per wave as the expansion does
        int qn = 0;
        for (uint64_t g = (uint64_t)wave * 64; g < pkl::kTileWords; g += (uint64_t)NW * 64) {
            const uint64_t wi = g + lane;
            uint64_t bits = (wi < pkl::kTileWords) ? ~a.visited[w0 + wi] : 0ull;
            const uint32_t c = (uint32_t)__popcll(bits);
            const uint32_t incl = wave_incl_scan(c, lane);
            const int total = (int)__shfl(incl, 63, 64);
            int off = qn + (int)(incl - c);
            int end_all = qn + total;
            const uint32_t base = (uint32_t)(wi * 64);
            for (;;) {
                while (bits && off < PKL_QCAP) {
                    q[off++] = base + (uint32_t)(__ffsll((unsigned long long)bits) - 1);
                    bits &= bits - 1;
                }
                if (end_all < PKL_QCAP) {
                    qn = end_all;
                    break;
                }
                wave_sync();
                run_queue((uint32_t)PKL_QCAP);
                wave_sync();
                off -= PKL_QCAP;
                end_all -= PKL_QCAP;
                qn = 0;
                if (end_all == 0) break;
            }
        }
        wave_sync();
        run_queue((uint32_t)qn);
        wave_sync();
        if ((uint32_t)lane < (uint32_t)qbn) probe(qb[lane]);
        if (any) s_any = 1;
        __syncthreads();
        uint32_t* dst = nxt + w0 * 2;
        for (uint32_t i = tid; i < WIN; i += TPB) {
            const uint32_t v = win[i];
            if (v) dst[i] = v;
        }
        if (tid == 0 && s_any) a.wflags[xp][T] = 1;
        __syncthreads();
    }
    block_add3<TPB>(words, probes, 0ull, &st->pull_words, &st->pull_probes, nullptr);
}

// Bookkeeping between two launches of k_pkl_step (one block): with advance,
// the layer end of every batched search (layer_end_body); then, unless the
// search is done, the list of tiles the next launch visits -- those that a
// list or their own expansion reached, or whose old words hold states to
// clear -- and a zero frontier count for every other tile.  The list is
// ordered by the tile's count in the layer just finalized, largest first (a
// counting sort on 256 buckets of the largest count): a tile's work in the
// next launch follows that count, and taking the longest tiles first
// shortens the launch's tail.
constexpr uint32_t PKL_MAX_TILES = 1320;  // 12!/9!
static_assert(pkl::Geo<12>::TILES == PKL_MAX_TILES, "lists cover at most 12 pancakes");
template <int TPB>
__global__ __launch_bounds__(TPB) void k_pkl_layer_end(LayerEndArgs a, StepArgs s, int advance) {
    constexpr int NW = TPB / 64;
    BfsDevState* st = a.st;
    if (advance) layer_end_body<TPB>(a);
    __syncthreads();
    if (__hip_atomic_load(&st->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    const uint32_t L = __hip_atomic_load(&st->layer, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (threadIdx.x == 0) {
        // split the next layer (finalize, then choose the direction of its
        // expansion) once the frontier could outgrow the unvisited states: the
        // next layer is at most a few times the current one
        const unsigned long long cnt = advance ? a.layer_counts[L] : 0ull;
        const unsigned long long vis = st->visited;
        const unsigned long long unv = a.nbits > vis ? a.nbits - vis : 0ull;
        st->split = (advance && st->direction == 1u && 4ull * cnt > unv) ? 1u : 0u;
    }
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    const int gp = (int)(L & 1u);
    const uint8_t* xf = s.xflags[gp];
    const uint8_t* wf = s.wflags[gp];
    const uint32_t* tc_cur = s.tcount + (uint64_t)(L % 3) * s.tiles;
    const uint32_t* tc_old = s.tcount + (uint64_t)((L + 2) % 3) * s.tiles;
    uint32_t* tc_new = s.tcount + (uint64_t)((L + 1) % 3) * s.tiles;
    const uint32_t tiles = s.tiles < PKL_MAX_TILES ? s.tiles : PKL_MAX_TILES;  // the host checks tiles <= 1320
    __shared__ uint16_t s_key[PKL_MAX_TILES];  // 0xffff: not visited, else 255 - bucket
    __shared__ uint32_t s_start[256];
    __shared__ uint32_t s_max;
    __shared__ uint32_t s_wsum[NW];
    for (uint32_t i = tid; i < 256; i += TPB) s_start[i] = 0;
    if (tid == 0) s_max = 0;
    __syncthreads();
    for (uint32_t T = tid; T < tiles; T += TPB) {
        const bool act = (xf[T] | wf[T]) != 0 || tc_old[T] != 0;
        if (act) {
            atomicMax(&s_max, tc_cur[T]);
        } else {
            tc_new[T] = 0;
        }
        s_key[T] = act ? 0u : 0xffffu;
    }
    __syncthreads();
    const uint32_t mx = s_max;
    for (uint32_t T = tid; T < tiles; T += TPB) {
        if (s_key[T] == 0xffffu) continue;
        const uint32_t b = (uint32_t)(((uint64_t)tc_cur[T] * 256u) / ((uint64_t)mx + 1u));  // 0..255
        s_key[T] = (uint16_t)(255u - b);
        atomicAdd(&s_start[255u - b], 1u);
    }
    __syncthreads();
    // exclusive scan of the bucket sizes (wave 0), largest counts first
    if (wave == 0) {
        uint32_t run = 0;
        for (uint32_t b0 = 0; b0 < 256; b0 += 64) {
            const uint32_t c = s_start[b0 + lane];
            const uint32_t incl = wave_incl_scan(c, lane);
            s_start[b0 + lane] = run + incl - c;
            run += __shfl(incl, 63, 64);
        }
        if (lane == 0) s_wsum[0] = run;
    }
    __syncthreads();
    for (uint32_t T = tid; T < tiles; T += TPB) {
        const uint32_t k = s_key[T];
        if (k == 0xffffu) continue;
        s.active[atomicAdd(&s_start[k], 1u)] = T;
    }
    if (tid == 0) st->active_tiles = s_wsum[0];
}

}  // namespace psk
