// Device-resident BFS kernels (three-bit and two-bit), templated on the puzzle
// functor of device_puzzles.h.  Included by the per-family translation units,
// which instantiate them; bfs.hip drives them.
//
// Three-bit layer L (Common/InMemoryThreeBitBFS.cpp:25-44) uses three bit
// arrays A[0..2]: cur = A[L%3], new = A[(L+1)%3], old = A[(L+2)%3].
//   k_bfs3_expand   scan cur -> expand every parent -> set children in new
//   k_bfs3_finalize new &= ~cur (odd cycles); new &= ~old; count; hash; old = 0
//   k_bfs3_layer_end bookkeeping: record the layer, rotate roles (layer += 1)
// The layer index lives in device memory, so the same launches (or a captured
// graph of them) serve every layer.
//
// Expand: the index space is cut into tiles of `tile_words` 64-bit words.  A
// workgroup takes a tile, keeps a bitmap of that tile's states in LDS, and
// expands every parent in it.  Children that land inside the tile are set with
// LDS atomics; the bitmap is OR-ed into HBM once at the end of the tile.  Other
// children go straight to HBM with a no-return atomicOr.  Parents are found
// with a per-wave compaction: each lane owns one word, popcounts are prefix
// summed across the wave, set-bit offsets go into a 256-entry LDS queue, and the
// queue is expanded 64 parents at a time so every lane does useful work.
#pragma once

#include "device_puzzles.h"
#include "ps_internal.h"

namespace psk {

constexpr int QCAP = 256;                 // per-wave compaction queue (entries)
constexpr uint32_t TOUCH_LDS_MAX = 65536;  // segments tracked in LDS per workgroup
constexpr uint64_t FIN_CHUNK_WORDS = 8192;  // words per finalize work item

struct BfsDevState {
    uint32_t layer;      // index of the current frontier layer
    uint32_t done;       // set when a layer came out empty or max_layers reached
    uint32_t error;      // bit0: a child index >= nbits was generated; bit1: a cross-tile list
                         // slot or fill count out of range; bit2: a list entry outside its tile
    uint32_t max_layers;
    unsigned long long tile_counter;
    unsigned long long chunk_counter;  // tiles taken by the first launch of a fused list search
    unsigned long long children;
    unsigned long long expanded;
    unsigned long long acc_count;
    unsigned long long acc_hash;
    unsigned long long layers_recorded;
    unsigned long long scan_words;       // cur words read by the expand kernels
    unsigned long long fin_full_words;   // words finalized (new/cur/old read, new/old written)
    unsigned long long fin_clear_words;  // words of old cleared without a finalize
    uint32_t pull;       // the current layer is expanded bottom-up (see pull_tile)
    uint32_t direction;  // 0: always top-down, 1: choose per layer, 2: always bottom-up
    uint32_t active_tiles;  // tiles the next fused list launch visits (StepArgs::active)
    uint32_t split;         // fused list search: the next launch only finalizes; the layer after it is
                            // then expanded top-down or bottom-up by the launches that follow
    unsigned long long visited;      // states in the layers recorded so far
    unsigned long long pull_words;   // words scanned by bottom-up layers (old and cur read)
    unsigned long long pull_probes;  // cur bits read outside the tile by bottom-up layers
};

struct ExpandArgs {
    psd::PuzzleParams prm;
    uint64_t* bits[3];
    uint64_t nbits;
    uint64_t nwords;
    uint64_t tile_words;
    uint64_t tile_begin, tile_end;
    uint64_t word_lo, word_hi;
    int32_t seg_shift;  // log2(bits per segment)
    uint32_t nseg;
    unsigned long long* seg_count;  // [3][nseg]
    uint32_t* touched;              // nseg bits (global)
    BfsDevState* st;
    int32_t persistent;
    int32_t use_window;
    uint8_t* xmark;  // byte per state for children outside the window; NULL: atomic OR into next
    int32_t skip_visited;  // skip cross-window children already in cur or old
    int32_t pull_capable;  // LDS holds a copy of the tile's cur words (bottom-up layers possible)
    const uint32_t* active;  // top-down layers visit st->active_tiles tiles listed here; NULL: every tile
};

// Cross-tile child lists of the 10..12-pancake search (bfs_pancake_lists.h).
struct ListArgs {
    uint32_t* lists;   // child ranks, [source tile][flip - 10][9!]
    uint32_t* counts;  // list fill counts, [source tile][slot of pancake_lists.h]
    uint8_t* flags;    // per tile: 1 when a list of this layer reaches it
    uint32_t tiles;    // tiles holding states: N!/9!
};

struct ApplyArgs {
    ListArgs x;
    uint64_t* bits[3];
    int32_t seg_shift;
    uint32_t nseg;
    uint32_t* touched;
    BfsDevState* st;
    int32_t persistent;
    uint64_t tile_begin, tile_end;
};

// One fused layer of the batched search with cross-tile lists (k_pkl_step).
struct StepArgs {
    uint64_t* bits[3];
    BfsDevState* st;
    uint32_t* lists[2];   // by parity of the launch that filled them
    uint32_t* counts[2];
    uint8_t* xflags[2];   // a list of this parity reaches the tile
    uint8_t* wflags[2];   // the tile's own expansion marked its next words
    uint32_t* tcount;     // [3][tiles] frontier states per tile, by layer mod 3
    uint32_t* active;     // [tiles] the tiles a launch visits (st->active_tiles of them)
    uint32_t tiles;
    uint32_t first_tile;  // the start state's tile (first launch)
    int32_t odd;
    int32_t hash;
    int32_t first;        // 1: expand the start layer only
    int32_t role;         // 0: finalize (and expand unless st->split); 1: expand only (after a split finalize)
    uint64_t* visited;    // every state of the layers finalized so far (NULL: top-down only)
    uint64_t nbits;
};

// bottom-up layers of the fused list search (k_pkl_pull)
struct PullArgs {
    uint64_t* bits[3];
    BfsDevState* st;
    const uint64_t* visited;
    uint8_t* wflags[2];
    uint32_t tiles;
    uint64_t nbits;
};

struct FinalizeArgs {
    uint64_t* bits[3];
    uint64_t nwords;
    uint64_t seg_words;      // words per segment (nwords when one segment)
    uint64_t chunks_per_seg;
    uint64_t chunk_begin, chunk_end;
    uint32_t nseg;
    int32_t odd;
    int32_t hash;
    int32_t persistent;
    unsigned long long* seg_count;
    uint32_t* touched;
    BfsDevState* st;
    uint8_t* xmark;  // folded into next and cleared (NULL when the expand kernel uses atomics)
};

struct LayerEndArgs {
    uint64_t nbits;
    uint32_t nseg;
    unsigned long long* seg_count;
    uint32_t* touched;
    unsigned long long* layer_counts;
    unsigned long long* layer_hashes;
    BfsDevState* st;
    int32_t bits;  // 3 or 2
    uint32_t* active;    // NULL, or the list of tiles the next top-down expand visits (ExpandArgs::active)
    uint64_t ntiles;
    uint64_t tile_bits;  // states per tile
    int32_t seg_shift;
    // k_bfs3_tile_step searches (four arrays, TileStepArgs): 1; touched_par as
    // in TileStepArgs::touched.  list_only: build the next launch's tile list
    // and nothing else (after the start layer's expansion).
    int32_t rot4;
    int32_t list_only;
    uint32_t* touched_par[2];
};

struct Bfs2Args {
    psd::PuzzleParams prm;
    uint64_t* arr;  // 2 bits per state, state i at bits 2(i%32).. of word i/32
    uint64_t nbits;
    uint64_t nwords;
    uint64_t tile_words;
    uint64_t ntiles;
    BfsDevState* st;
};

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__device__ __forceinline__ uint64_t word_hash(uint64_t j, uint64_t w) {
    return w ? splitmix64(splitmix64(j) ^ w) : 0ull;
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, int lane) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ void gor(uint32_t* p, uint32_t v) {
    __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void lor(uint32_t* p, uint32_t v) {
    __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Block-wide sums of up to three u64 counters, then one atomicAdd each (a
// NULL destination skips that counter).  Every thread of the block must call it.
template <int TPB>
__device__ __forceinline__ void block_add3(unsigned long long a, unsigned long long b, unsigned long long c,
                                           unsigned long long* ga, unsigned long long* gb,
                                           unsigned long long* gc) {
    __shared__ unsigned long long red[3][TPB / 64];
    __syncthreads();
    a = wave_sum(a);
    b = wave_sum(b);
    c = wave_sum(c);
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0) {
        red[0][wave] = a;
        red[1][wave] = b;
        red[2][wave] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long sa = 0, sb = 0, sc = 0;
        for (int w = 0; w < TPB / 64; ++w) {
            sa += red[0][w];
            sb += red[1][w];
            sc += red[2][w];
        }
        if (sa && ga) atomicAdd(ga, sa);
        if (sb && gb) atomicAdd(gb, sb);
        if (sc && gc) atomicAdd(gc, sc);
    }
}

// The 64 mark bytes (0 or 1) of one bit-array word as its bits, clearing
// them; bytes are read and written as four 16-byte vectors.
// puzzles that can name the slots whose child may leave the segment
// (PZ::cross_slots(index, seg_bits)); the others expand every slot
template <class PZ, class = void>
struct HasCrossSlots : std::false_type {};
template <class PZ>
struct HasCrossSlots<PZ, std::void_t<decltype(PZ::cross_slots(0ull, 0))>> : std::true_type {};

__device__ __forceinline__ uint64_t fold_marks(uint8_t* p) {
    uint4* v = reinterpret_cast<uint4*>(p);
    uint64_t bits = 0;
    bool any = false;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint4 x = v[k];
        const uint32_t u[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            // bytes b0..b3 in {0,1}: the product gathers them at bits 28..31
            const uint32_t nib = ((u[e] & 0x01010101u) * 0x10204080u) >> 28;
            bits |= (uint64_t)nib << (16 * k + 4 * e);
            any |= u[e] != 0u;
        }
    }
    if (any) {
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = make_uint4(0u, 0u, 0u, 0u);
    }
    return bits;
}

// LDS bytes the expand kernel needs besides its static shared variables.
inline size_t bfs3_lds_bytes(int tpb, uint64_t tile_words, int use_window, uint32_t nseg, int pull_capable = 0) {
    size_t b = (size_t)(tpb / 64) * QCAP * 4;
    if (use_window) b += (size_t)tile_words * 8;
    if (pull_capable) b += (size_t)tile_words * 8;
    if (nseg > 1 && nseg <= TOUCH_LDS_MAX) b += (size_t)((nseg + 31) / 32) * 4;
    return b;
}

// LDS of the expand kernels: the tile window, one compaction queue per wave,
// and the segment "touched" bits of the job being expanded.
struct ExpandLds {
    uint32_t* win;
    uint64_t* cwin;  // cur words of the tile (bottom-up layers)
    uint32_t* queues;
    uint32_t* tl;
    uint32_t tl_words;  // 0: touched bits go straight to HBM
};

struct ExpandTally {
    unsigned long long children = 0, parents = 0, scan = 0;
};

// The arrays one tile's expansion reads and writes.
struct TileRoles {
    const uint64_t* cur;               // the frontier expanded
    uint32_t* nxt;                     // the next layer's marks
    const uint64_t* old;               // the layer before cur (skip_visited)
    const unsigned long long* sc_cur;  // per-segment counts of cur; NULL: the caller knows the tile holds states
    uint32_t* touched;                 // segments that receive marks
};

// Roles of the three rotating arrays at layer L.
__device__ __forceinline__ TileRoles roles3(const ExpandArgs& a, uint32_t L) {
    TileRoles r;
    r.cur = a.bits[L % 3];
    r.nxt = reinterpret_cast<uint32_t*>(a.bits[(L + 1) % 3]);
    r.old = a.bits[(L + 2) % 3];
    r.sc_cur = a.seg_count + (uint64_t)(L % 3) * a.nseg;
    r.touched = a.touched;
    return r;
}

// Expands every parent of one tile of one search.  All threads of the block
// call it with the same tile; it starts and ends with a block barrier.
template <class PZ, int TPB>
__device__ __forceinline__ void expand_tile_roles(const ExpandArgs& a, const TileRoles& rl, uint64_t tile,
                                                  const ExpandLds& lds, uint32_t* s_local, ExpandTally& tally) {
    BfsDevState* st = a.st;
    const uint64_t* __restrict__ cur = rl.cur;
    uint32_t* nxt = rl.nxt;
    const uint64_t* __restrict__ old = rl.old;
    const unsigned long long* sc_cur = rl.sc_cur;
    uint32_t* touched = rl.touched;
    constexpr int NW = TPB / 64;
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    uint32_t* win = lds.win;
    uint32_t* q = lds.queues + wave * QCAP;
    const bool track = a.nseg > 1;

    const uint64_t w0 = tile * a.tile_words;
    const uint64_t w1 = (w0 + a.tile_words < a.nwords) ? w0 + a.tile_words : a.nwords;
    const uint64_t pw0 = w0 > a.word_lo ? w0 : a.word_lo;
    const uint64_t pw1 = w1 < a.word_hi ? w1 : a.word_hi;
    bool live = pw0 < pw1;
    if (live && track && sc_cur) {
        const uint64_t s0 = (pw0 * 64) >> a.seg_shift, s1 = (pw1 * 64 - 1) >> a.seg_shift;
        live = false;
        for (uint64_t s = s0; s <= s1; ++s) live |= sc_cur[s] != 0;
    }
    if (!live) {
        __syncthreads();
        return;
    }
    if (tid == 0) *s_local = 0;
    const uint32_t nwin = (uint32_t)(w1 - w0) * 2;
    if (a.use_window)
        for (uint32_t i = tid; i < nwin; i += TPB) win[i] = 0;
    __syncthreads();

    if (tid == 0) tally.scan += pw1 - pw0;
    const uint64_t win_lo = w0 * 64, win_hi = w1 * 64;
    // states of the window below nbits: one unsigned compare places a child
    // in the window; a child at or past nbits takes the other path and is reported
    const uint64_t win_states = a.use_window ? (win_hi < a.nbits ? win_hi : a.nbits) - win_lo : 0;
    bool local_any = false;
    uint32_t n_local = 0;  // children marked in the window (added to the tally at the end)
    auto mark = [&](uint64_t child) {
        const uint64_t rel = child - win_lo;
        if (rel < win_states) {
            lor(&win[rel >> 5], 1u << (rel & 31));
            local_any = true;
            ++n_local;
            return;
        }
        if (child >= a.nbits) {
            atomicOr(&st->error, 1u);
            return;
        }
        ++tally.children;
        {
            if (a.skip_visited && (((cur[child >> 6] | old[child >> 6]) >> (child & 63)) & 1ull)) return;
            if (a.xmark) a.xmark[child] = 1;
            else gor(&nxt[child >> 5], 1u << (child & 31));
            if (track) {
                const uint64_t s = child >> a.seg_shift;
                if (lds.tl_words) lor(&lds.tl[s >> 5], 1u << (s & 31));
                else gor(&touched[s >> 5], 1u << (s & 31));
            }
        }
    };
    auto run_batch = [&](uint32_t n) {
        for (uint32_t j = lane; j < n; j += 64) {
            const uint64_t parent = win_lo + q[j];
            ++tally.parents;
            PZ::expand(parent, 0u, a.prm, [&](int, uint64_t child, uint32_t) { mark(child); });
        }
    };

    int qn = 0;  // wave-uniform queue fill
    for (uint64_t g = pw0 + (uint64_t)wave * 64; g < pw1; g += (uint64_t)NW * 64) {
        const uint64_t wi = g + lane;
        uint64_t bits = (wi < pw1) ? cur[wi] : 0ull;
        const uint32_t c = (uint32_t)__popcll(bits);
        const uint32_t incl = wave_incl_scan(c, lane);
        const int total = (int)__shfl(incl, 63, 64);
        int off = qn + (int)(incl - c);
        int end_all = qn + total;
        const uint32_t base = (uint32_t)((wi - w0) * 64);
        for (;;) {
            while (bits && off < QCAP) {
                q[off++] = base + (uint32_t)(__ffsll((unsigned long long)bits) - 1);
                bits &= bits - 1;
            }
            if (end_all < QCAP) {
                qn = end_all;
                break;
            }
            wave_sync();
            run_batch(QCAP);
            wave_sync();
            off -= QCAP;
            end_all -= QCAP;
            qn = 0;
            if (end_all == 0) break;
        }
    }
    wave_sync();
    run_batch((uint32_t)qn);
    tally.children += n_local;
    if (local_any) *s_local = 1;
    __syncthreads();
    if (a.use_window) {
        // with byte marks nothing else writes this range of next, so a tile
        // that owns its whole range (processed once per layer) stores plainly
        uint32_t* dst = nxt + w0 * 2;
        const bool own = a.xmark && pw0 == w0 && pw1 == w1;
        for (uint32_t i = tid; i < nwin; i += TPB) {
            const uint32_t v = win[i];
            if (v) {
                if (own) dst[i] = v;
                else gor(&dst[i], v);
            }
        }
    }
    if (tid == 0 && track && *s_local) {
        const uint64_t s0 = win_lo >> a.seg_shift, s1 = (win_hi - 1) >> a.seg_shift;
        for (uint64_t s = s0; s <= s1; ++s) {
            if (lds.tl_words) lor(&lds.tl[s >> 5], 1u << (s & 31));
            else gor(&touched[s >> 5], 1u << (s & 31));
        }
    }
    __syncthreads();
}

template <class PZ, int TPB>
__device__ __forceinline__ void expand_tile(const ExpandArgs& a, uint64_t tile, const ExpandLds& lds,
                                            uint32_t* s_local, ExpandTally& tally) {
    expand_tile_roles<PZ, TPB>(a, roles3(a, a.st->layer), tile, lds, s_local, tally);
}

// Bottom-up expansion of one tile (direction-optimizing BFS): every unvisited
// state of the tile (not in old or cur) looks for a parent among its own
// children, which are its neighbours because every move here is reversible.
// In-tile neighbours are tested in an LDS copy of the tile's cur words; a
// neighbour outside the tile costs one read of cur, issued only while no
// parent has been found.  The tile's new bits are written with plain stores:
// no other tile writes them.  Requires use_window and pull_capable.
template <class PZ, int TPB>
__device__ __forceinline__ void pull_tile(const ExpandArgs& a, uint64_t tile, const ExpandLds& lds,
                                          uint32_t* s_local, ExpandTally& tally, unsigned long long& probes) {
    BfsDevState* st = a.st;
    const uint32_t L = st->layer;
    const uint64_t* __restrict__ cur = a.bits[L % 3];
    uint32_t* nxt = reinterpret_cast<uint32_t*>(a.bits[(L + 1) % 3]);
    const uint64_t* __restrict__ old = a.bits[(L + 2) % 3];
    constexpr int NW = TPB / 64;
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    uint32_t* win = lds.win;
    uint64_t* cw = lds.cwin;
    uint32_t* q = lds.queues + wave * QCAP;
    const bool track = a.nseg > 1;

    const uint64_t w0 = tile * a.tile_words;
    const uint64_t w1 = (w0 + a.tile_words < a.nwords) ? w0 + a.tile_words : a.nwords;
    const uint64_t pw0 = w0 > a.word_lo ? w0 : a.word_lo;
    const uint64_t pw1 = w1 < a.word_hi ? w1 : a.word_hi;
    if (pw0 >= pw1) {
        __syncthreads();
        return;
    }
    if (tid == 0) *s_local = 0;
    const uint32_t nw = (uint32_t)(w1 - w0);
    for (uint32_t i = tid; i < nw * 2; i += TPB) win[i] = 0;
    for (uint32_t i = tid; i < nw; i += TPB) cw[i] = cur[w0 + i];
    __syncthreads();
    if (tid == 0) tally.scan += pw1 - pw0;
    const uint64_t win_lo = w0 * 64, win_hi = w1 * 64;
    bool local_any = false;
    auto run_batch = [&](uint32_t n) {
        for (uint32_t j = lane; j < n; j += 64) {
            const uint32_t off = q[j];
            bool hit = false;
            PZ::expand(win_lo + off, 0u, a.prm, [&](int, uint64_t nb, uint32_t) {
                if (nb >= win_lo && nb < win_hi) {
                    const uint64_t o = nb - win_lo;
                    hit |= ((cw[o >> 6] >> (o & 63)) & 1ull) != 0ull;
                } else if (!hit && nb < a.nbits) {
                    ++probes;
                    hit = ((cur[nb >> 6] >> (nb & 63)) & 1ull) != 0ull;
                }
            });
            if (hit) {
                lor(&win[off >> 5], 1u << (off & 31));
                local_any = true;
            }
        }
    };
    int qn = 0;
    for (uint64_t g = pw0 + (uint64_t)wave * 64; g < pw1; g += (uint64_t)NW * 64) {
        const uint64_t wi = g + lane;
        uint64_t bits = 0;
        if (wi < pw1) {
            bits = ~(old[wi] | cw[wi - w0]);
            const uint64_t lo = wi * 64;
            if (lo + 64 > a.nbits) bits &= (a.nbits > lo) ? ((1ull << (a.nbits - lo)) - 1ull) : 0ull;
        }
        const uint32_t c = (uint32_t)__popcll(bits);
        const uint32_t incl = wave_incl_scan(c, lane);
        const int total = (int)__shfl(incl, 63, 64);
        int off = qn + (int)(incl - c);
        int end_all = qn + total;
        const uint32_t base = (uint32_t)((wi - w0) * 64);
        for (;;) {
            while (bits && off < QCAP) {
                q[off++] = base + (uint32_t)(__ffsll((unsigned long long)bits) - 1);
                bits &= bits - 1;
            }
            if (end_all < QCAP) {
                qn = end_all;
                break;
            }
            wave_sync();
            run_batch(QCAP);
            wave_sync();
            off -= QCAP;
            end_all -= QCAP;
            qn = 0;
            if (end_all == 0) break;
        }
    }
    wave_sync();
    run_batch((uint32_t)qn);
    if (local_any) *s_local = 1;
    __syncthreads();
    // only words of the tile's own state range [pw0, pw1) can hold new bits
    uint32_t* dst = nxt + w0 * 2;
    for (uint32_t i = (uint32_t)(pw0 - w0) * 2 + tid; i < (uint32_t)(pw1 - w0) * 2; i += TPB) {
        const uint32_t v = win[i];
        if (v) dst[i] = v;
    }
    if (tid == 0 && track && *s_local) {
        const uint64_t s0 = (pw0 * 64) >> a.seg_shift, s1 = (pw1 * 64 - 1) >> a.seg_shift;
        for (uint64_t s = s0; s <= s1; ++s) {
            if (lds.tl_words) lor(&lds.tl[s >> 5], 1u << (s & 31));
            else gor(&a.touched[s >> 5], 1u << (s & 31));
        }
    }
    __syncthreads();
}

template <class PZ, int TPB>
__global__ __launch_bounds__(TPB) void k_bfs3_expand(ExpandArgs a) {
    extern __shared__ uint32_t smem[];
    constexpr int NW = TPB / 64;
    BfsDevState* st = a.st;
    if (st->done) return;
    const int tid = threadIdx.x;
    const uint32_t win_u32 = a.use_window ? (uint32_t)(a.tile_words * 2) : 0u;
    const uint32_t cwin_u32 = a.pull_capable ? (uint32_t)(a.tile_words * 2) : 0u;
    ExpandLds lds;
    lds.win = smem;
    lds.cwin = reinterpret_cast<uint64_t*>(smem + win_u32);
    lds.queues = smem + win_u32 + cwin_u32;
    lds.tl = lds.queues + NW * QCAP;
    lds.tl_words = (a.nseg > 1 && a.nseg <= TOUCH_LDS_MAX) ? (a.nseg + 31) / 32 : 0u;
    const bool pull = a.pull_capable && st->pull;
    unsigned long long probes = 0;
    __shared__ unsigned long long s_tile;
    __shared__ uint32_t s_local;
    for (uint32_t i = tid; i < lds.tl_words; i += TPB) lds.tl[i] = 0;

    ExpandTally tally;
    uint64_t iter = 0;
    // a top-down layer with a tile list visits only the tiles holding frontier
    // states: a sparse layer costs no pass over every tile and no contended
    // counter per empty tile
    const bool listed = a.active != nullptr && !pull;
    for (;;) {
        if (tid == 0) {
            if (listed) {
                const unsigned long long k = atomicAdd(&st->tile_counter, 1ull);
                s_tile = k < st->active_tiles ? (uint64_t)a.active[k] : a.tile_end;
            } else if (a.persistent) {
                s_tile = a.tile_begin + atomicAdd(&st->tile_counter, 1ull);
            } else {
                s_tile = (iter == 0) ? a.tile_begin + blockIdx.x : ~0ull;
            }
        }
        __syncthreads();
        const uint64_t tile = s_tile;
        ++iter;
        if (tile >= a.tile_end) break;
        if (pull) pull_tile<PZ, TPB>(a, tile, lds, &s_local, tally, probes);
        else expand_tile<PZ, TPB>(a, tile, lds, &s_local, tally);
    }
    __syncthreads();
    for (uint32_t i = tid; i < lds.tl_words; i += TPB) {
        const uint32_t v = lds.tl[i];
        if (v) gor(&a.touched[i], v);
    }
    if (pull) block_add3<TPB>(tally.scan, probes, 0ull, &st->pull_words, &st->pull_probes, nullptr);
    else block_add3<TPB>(tally.children, tally.parents, tally.scan, &st->children, &st->expanded, &st->scan_words);
}

// Finalize of one chunk of one search: new &= ~cur [odd]; new &= ~old; per-segment
// and layer counts; hash; old = 0.  Returns the chunk's (count, hash, full, cleared)
// through the tally; every thread of the block calls it.
struct FinTally {
    unsigned long long cnt = 0, hash = 0, full = 0, clear = 0;
};

// One word of the finalize: new &= ~old [& ~cur]; returns the new word.
__device__ __forceinline__ uint64_t fin_word(uint64_t v, uint64_t o, uint64_t c, bool odd) {
    v &= ~o;
    if (odd) v &= ~c;
    return v;
}

// Words [w0, w1) of a touched segment, two words per thread and step, the
// loads of FIN_UNROLL steps issued before any store (the arrays are distinct
// allocations).  XM: fold the cross-window mark bytes into new.
constexpr int FIN_UNROLL = 4;

template <int TPB, bool XM, int UN = FIN_UNROLL>
__device__ __forceinline__ void finalize_words(const uint64_t* __restrict__ cur, uint64_t* __restrict__ nw,
                                               uint64_t* __restrict__ old, uint8_t* __restrict__ xmark, uint64_t w0,
                                               uint64_t w1, bool odd, bool hash, unsigned long long& cnt,
                                               unsigned long long& h) {
    constexpr int U = XM ? 2 : UN;
    const uint64_t npair = (w1 - w0) / 2;  // w0 is even: 16-byte accesses are aligned
    for (uint64_t p0 = threadIdx.x; p0 < npair; p0 += (uint64_t)TPB * U) {
        ulonglong2 vn[U], vo[U], vc[U];
        uint64_t xb[U][2];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t p = p0 + (uint64_t)u * TPB;
            if (p < npair) {
                const uint64_t j = w0 + 2 * p;
                vn[u] = *reinterpret_cast<const ulonglong2*>(nw + j);
                vo[u] = *reinterpret_cast<const ulonglong2*>(old + j);
                vc[u] = odd ? *reinterpret_cast<const ulonglong2*>(cur + j) : make_ulonglong2(0ull, 0ull);
                if (XM) {
                    xb[u][0] = fold_marks(xmark + j * 64);
                    xb[u][1] = fold_marks(xmark + (j + 1) * 64);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t p = p0 + (uint64_t)u * TPB;
            if (p < npair) {
                const uint64_t j = w0 + 2 * p;
                uint64_t a = vn[u].x, b = vn[u].y;
                if (XM) {
                    a |= xb[u][0];
                    b |= xb[u][1];
                }
                a = fin_word(a, vo[u].x, vc[u].x, odd);
                b = fin_word(b, vo[u].y, vc[u].y, odd);
                // frontiers are sparse: store only the words that change
                if (a != vn[u].x || b != vn[u].y) *reinterpret_cast<ulonglong2*>(nw + j) = make_ulonglong2(a, b);
                if (vo[u].x | vo[u].y) *reinterpret_cast<ulonglong2*>(old + j) = make_ulonglong2(0ull, 0ull);
                cnt += (unsigned long long)(__popcll(a) + __popcll(b));
                if (hash) h += word_hash(j, a) + word_hash(j + 1, b);
            }
        }
    }
    // an odd word count leaves one word
    if (((w1 - w0) & 1) && threadIdx.x == 0) {
        const uint64_t j = w1 - 1;
        const uint64_t n0 = nw[j], o = old[j];
        uint64_t v = n0;
        if (XM) v |= fold_marks(xmark + j * 64);
        v = fin_word(v, o, odd ? cur[j] : 0ull, odd);
        if (v != n0) nw[j] = v;
        if (o) old[j] = 0;
        cnt += (unsigned long long)__popcll(v);
        if (hash) h += word_hash(j, v);
    }
}

// Finalize of one chunk of one search: new &= ~cur [odd]; new &= ~old; per-segment
// and layer counts; hash; old = 0.  Every thread of the block calls it.
template <int TPB>
__device__ __forceinline__ void finalize_chunk(const FinalizeArgs& a, uint64_t c, FinTally& tally) {
    BfsDevState* st = a.st;
    const uint32_t L = st->layer;
    const uint64_t* cur = a.bits[L % 3];
    uint64_t* nw = a.bits[(L + 1) % 3];
    uint64_t* old = a.bits[(L + 2) % 3];
    unsigned long long* sc_new = a.seg_count + (uint64_t)((L + 1) % 3) * a.nseg;
    const unsigned long long* sc_old = a.seg_count + (uint64_t)((L + 2) % 3) * a.nseg;
    const bool track = a.nseg > 1;
    __shared__ unsigned long long red[2][TPB / 64];
    const uint64_t s = c / a.chunks_per_seg;
    const uint64_t w0 = s * a.seg_words + (c % a.chunks_per_seg) * FIN_CHUNK_WORDS;
    uint64_t w1 = w0 + FIN_CHUNK_WORDS;
    const uint64_t seg_end = (s + 1) * a.seg_words;
    if (w1 > seg_end) w1 = seg_end;
    if (w1 > a.nwords) w1 = a.nwords;
    if (w0 >= w1) return;
    const bool touched = !track || ((a.touched[s >> 5] >> (s & 31)) & 1u);
    const bool old_live = !track || sc_old[s] != 0;
    unsigned long long cnt = 0, h = 0;
    if (touched) {
        const bool odd = a.odd != 0, hash = a.hash != 0;
        if (a.xmark && !st->pull) finalize_words<TPB, true>(cur, nw, old, a.xmark, w0, w1, odd, hash, cnt, h);
        else finalize_words<TPB, false>(cur, nw, old, nullptr, w0, w1, odd, hash, cnt, h);
    } else if (old_live) {
        for (uint64_t j = w0 + threadIdx.x; j < w1; j += TPB) old[j] = 0;
    }
    if (threadIdx.x == 0) {
        if (touched) tally.full += w1 - w0;
        else if (old_live) tally.clear += w1 - w0;
    }
    if (touched) {
        cnt = wave_sum(cnt);
        h = wave_sum(h);
        const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
        if (lane == 0) {
            red[0][wave] = cnt;
            red[1][wave] = h;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long sc = 0, sh = 0;
            for (int w = 0; w < TPB / 64; ++w) {
                sc += red[0][w];
                sh += red[1][w];
            }
            if (sc) atomicAdd(&sc_new[s], sc);
            tally.cnt += sc;
            tally.hash += sh;
        }
        __syncthreads();
    }
}

__device__ __forceinline__ void flush_fin_tally(BfsDevState* st, const FinTally& t) {
    if (threadIdx.x == 0) {
        if (t.cnt) atomicAdd(&st->acc_count, t.cnt);
        if (t.hash) atomicAdd(&st->acc_hash, t.hash);
        if (t.full) atomicAdd(&st->fin_full_words, t.full);
        if (t.clear) atomicAdd(&st->fin_clear_words, t.clear);
    }
}

template <int TPB>
__global__ __launch_bounds__(TPB) void k_bfs3_finalize(FinalizeArgs a) {
    BfsDevState* st = a.st;
    if (st->done) return;
    __shared__ unsigned long long s_chunk;
    FinTally tally;
    uint64_t iter = 0;
    for (;;) {
        if (threadIdx.x == 0) {
            // chunks carry similar work: a static stride, no shared counter
            if (a.persistent) s_chunk = a.chunk_begin + blockIdx.x + iter * gridDim.x;
            else s_chunk = (iter == 0) ? a.chunk_begin + blockIdx.x : ~0ull;
        }
        __syncthreads();
        const uint64_t c = s_chunk;
        ++iter;
        __syncthreads();
        if (c >= a.chunk_end) break;
        finalize_chunk<TPB>(a, c, tally);
    }
    flush_fin_tally(st, tally);
}

// One layer of a batched three-bit search whose children outside the LDS
// window are OR'ed into the next array (PS_CROSS_ATOMIC), with the finalize
// and the expansion of each tile in one launch.  Four arrays rotate: the
// frontier F_j and the marks M_j of layer j live in bits[j % 4].  The launch
// at layer L finalizes M_{L+1} in place into F_{L+1} (new &= ~F_L [odd],
// new &= ~F_{L-1}; F_{L-1}'s words cleared) and expands F_{L+1} into
// bits[(L+2) % 4], which the launch before cleared, so no expansion writes an
// array that another tile's finalize reads.  touched[j % 2] flags the
// segments holding marks of layer j, seg_count[j % 4] holds F_j's counts.
// Each tile lies inside one segment (or there is one segment); the tile list
// (ExpandArgs::active) names every tile with marks or F_{L-1} states.
struct TileStepArgs {
    uint64_t* bits[4];
    unsigned long long* seg_count;  // [4][nseg]
    uint32_t* touched[2];
    uint64_t seg_words;
    int32_t odd;
    int32_t hash;
};

template <class PZ, int TPB>
__global__ __launch_bounds__(TPB) void k_bfs3_tile_step(ExpandArgs a, TileStepArgs t) {
    extern __shared__ uint32_t smem[];
    constexpr int NW = TPB / 64;
    BfsDevState* st = a.st;
    if (st->done) return;
    const int tid = threadIdx.x;
    const uint32_t L = st->layer;
    uint64_t* mk = t.bits[(L + 1) & 3u];        // M_{L+1}, finalized into F_{L+1}
    const uint64_t* cur = t.bits[L & 3u];       // F_L
    uint64_t* old = t.bits[(L + 3) & 3u];       // F_{L-1}, cleared
    unsigned long long* sc_new = t.seg_count + (uint64_t)((L + 1) & 3u) * a.nseg;
    const unsigned long long* sc_old = t.seg_count + (uint64_t)((L + 3) & 3u) * a.nseg;
    const uint32_t* tin = t.touched[(L + 1) & 1u];
    TileRoles rl;
    rl.cur = mk;
    rl.nxt = reinterpret_cast<uint32_t*>(t.bits[(L + 2) & 3u]);
    rl.old = nullptr;
    rl.sc_cur = nullptr;
    rl.touched = t.touched[L & 1u];
    const bool track = a.nseg > 1;
    const bool odd = t.odd != 0, hash = t.hash != 0;
    ExpandLds lds;
    lds.win = smem;
    lds.cwin = nullptr;
    lds.queues = smem + (uint32_t)(a.tile_words * 2);
    lds.tl = lds.queues + NW * QCAP;
    lds.tl_words = (track && a.nseg <= TOUCH_LDS_MAX) ? (a.nseg + 31) / 32 : 0u;
    __shared__ unsigned long long s_tile, s_cnt;
    __shared__ unsigned long long red[2][NW];
    __shared__ uint32_t s_local;
    for (uint32_t i = tid; i < lds.tl_words; i += TPB) lds.tl[i] = 0;

    ExpandTally tally;
    FinTally ft;
    for (;;) {
        if (tid == 0) {
            const unsigned long long k = atomicAdd(&st->tile_counter, 1ull);
            s_tile = k < st->active_tiles ? (uint64_t)a.active[k] : a.tile_end;
        }
        __syncthreads();
        const uint64_t tile = s_tile;
        if (tile >= a.tile_end) break;
        const uint64_t w0 = tile * a.tile_words;
        const uint64_t w1 = (w0 + a.tile_words < a.nwords) ? w0 + a.tile_words : a.nwords;
        const uint64_t sg = track ? w0 / t.seg_words : 0;
        const bool touched = !track || ((tin[sg >> 5] >> (sg & 31)) & 1u);
        const bool old_live = !track || sc_old[sg] != 0;
        if (touched) {
            unsigned long long cnt = 0, h = 0;
            // two steps in flight: the expansion's registers bound the kernel
            finalize_words<TPB, false, 2>(cur, mk, old, nullptr, w0, w1, odd, hash, cnt, h);
            cnt = wave_sum(cnt);
            h = wave_sum(h);
            const int wave = tid / 64, lane = tid % 64;
            if (lane == 0) {
                red[0][wave] = cnt;
                red[1][wave] = h;
            }
            __syncthreads();
            if (tid == 0) {
                unsigned long long sc = 0, sh = 0;
                for (int w = 0; w < NW; ++w) {
                    sc += red[0][w];
                    sh += red[1][w];
                }
                if (sc) atomicAdd(&sc_new[sg], sc);
                ft.cnt += sc;
                ft.hash += sh;
                ft.full += w1 - w0;
                s_cnt = sc;
            }
        } else {
            if (old_live)
                for (uint64_t j = w0 + tid; j < w1; j += TPB) old[j] = 0;
            if (tid == 0) {
                if (old_live) ft.clear += w1 - w0;
                s_cnt = 0;
            }
        }
        // the tile's new frontier words, written above by this block, are read
        // by the expansion after the barrier
        __syncthreads();
        if (s_cnt) expand_tile_roles<PZ, TPB>(a, rl, tile, lds, &s_local, tally);
        else __syncthreads();
    }
    __syncthreads();
    for (uint32_t i = tid; i < lds.tl_words; i += TPB) {
        const uint32_t v = lds.tl[i];
        if (v) gor(&rl.touched[i], v);
    }
    block_add3<TPB>(tally.children, tally.parents, tally.scan, &st->children, &st->expanded, &st->scan_words);
    flush_fin_tally(st, ft);
}

// Bookkeeping at the end of a layer, run by one block: record the layer, rotate
// the roles, reset the per-layer counters.
template <int TPB>
__device__ __forceinline__ void layer_end_body(const LayerEndArgs& a) {
    BfsDevState* st = a.st;
    if (st->done) return;
    const uint32_t L = st->layer;
    __shared__ uint32_t s_list;
    if (a.rot4 && !a.list_only) {
        // F_{L-1} was cleared by the launch just ended: its counts next serve
        // F_{L+3}; the marks of layer L+1 are finalized
        unsigned long long* sc_free = a.seg_count + (uint64_t)((L + 3) & 3u) * a.nseg;
        for (uint32_t i = threadIdx.x; i < a.nseg; i += TPB) sc_free[i] = 0;
        if (a.nseg > 1) {
            uint32_t* t = a.touched_par[(L + 1) & 1u];
            for (uint32_t i = threadIdx.x; i < (a.nseg + 31) / 32; i += TPB) t[i] = 0;
        }
    } else if (a.bits == 3 && !a.rot4) {
        unsigned long long* sc_next_new = a.seg_count + (uint64_t)((L + 2) % 3) * a.nseg;
        for (uint32_t i = threadIdx.x; i < a.nseg; i += TPB) sc_next_new[i] = 0;
        if (a.nseg > 1)
            for (uint32_t i = threadIdx.x; i < (a.nseg + 31) / 32; i += TPB) a.touched[i] = 0;
    }
    __syncthreads();
    if (threadIdx.x == 0 && !a.list_only) {
        const unsigned long long cnt = st->acc_count;
        st->visited += cnt;
        // the next layer goes bottom-up when its frontier outnumbers the states
        // not visited yet (which bound the bottom-up work)
        const unsigned long long unvisited = a.nbits > st->visited ? a.nbits - st->visited : 0ull;
        st->pull = st->direction == 2 || (st->direction == 1 && cnt > unvisited);
        if (cnt == 0) {
            st->done = 1;
        } else {
            a.layer_counts[L + 1] = cnt;
            a.layer_hashes[L + 1] = st->acc_hash;
            st->layers_recorded = L + 2;
            st->layer = L + 1;
            if (L + 2 >= st->max_layers) st->done = 1;
        }
        st->acc_count = 0;
        st->acc_hash = 0;
    }
    if (threadIdx.x == 0) {
        st->tile_counter = 0;
        st->chunk_counter = 0;
        s_list = (a.active && !st->done && !st->pull) ? 1u : 0u;
    }
    __syncthreads();
    if (!s_list) return;
    // the tiles the next (top-down) expand visits: those overlapping a segment
    // with states in the layer just finalized, in tile order
    constexpr int NW = TPB / 64;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const unsigned long long* sc = a.seg_count + (uint64_t)((L + 1) % 3) * a.nseg;
    // a four-array launch at layer Ln finalizes the segments with marks of
    // layer Ln+1 and clears those holding F_{Ln-1}
    const uint32_t Ln = a.list_only ? L : L + 1;
    const uint32_t* tmark = a.rot4 ? a.touched_par[(Ln + 1) & 1u] : nullptr;
    if (a.rot4) sc = a.seg_count + (uint64_t)((Ln + 3) & 3u) * a.nseg;
    __shared__ uint32_t s_wsum[NW];
    uint32_t base = 0;
    for (uint64_t t0 = 0; t0 < a.ntiles; t0 += TPB) {
        const uint64_t T = t0 + threadIdx.x;
        bool live = false;
        if (T < a.ntiles) {
            if (a.nseg <= 1) {
                live = true;
            } else {
                const uint64_t b0 = T * a.tile_bits;
                const uint64_t b1 = (b0 + a.tile_bits < a.nbits ? b0 + a.tile_bits : a.nbits) - 1;
                const uint64_t s0 = b0 >> a.seg_shift, s1 = b1 >> a.seg_shift;
                if (s1 - s0 >= 64) live = true;  // many small segments: visit the tile
                else
                    for (uint64_t sg = s0; sg <= s1 && !live; ++sg)
                        live = sc[sg] != 0 || (tmark && ((tmark[sg >> 5] >> (sg & 31)) & 1u));
            }
        }
        const unsigned long long m = __ballot(live);
        if (lane == 0) s_wsum[wave] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t off = base, total = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const uint32_t c = s_wsum[w];
            off += (w < wave) ? c : 0u;
            total += c;
        }
        if (live) a.active[off + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = (uint32_t)T;
        base += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) st->active_tiles = base;
}

template <int TPB>
__global__ __launch_bounds__(TPB) void k_bfs_layer_end(LayerEndArgs a) {
    layer_end_body<TPB>(a);
}

// ---------------------------------------------------------------- two-bit

// Common/InMemoryTwoBitBFS.cpp:52-69: UNVISITED=0, OLD=1, CUR/NEXT = 2/3 swapped
// every layer.  One persistent pass per layer: CUR fields of a word become OLD
// (atomic xor, other lanes may be setting NEXT in the same word), their states
// are expanded, and each child that is UNVISITED is moved to NEXT with a
// compare-and-swap that also counts it.
template <class PZ, int TPB>
__global__ __launch_bounds__(TPB) void k_bfs2_layer(Bfs2Args a) {
    extern __shared__ uint32_t smem[];
    BfsDevState* st = a.st;
    if (st->done) return;
    const uint32_t L = st->layer;
    const uint32_t CUR = (L & 1) ? 3u : 2u, NEXT = (L & 1) ? 2u : 3u, OLD = 1u;
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    constexpr int NW = TPB / 64;
    uint32_t* q = smem + wave * QCAP;
    uint32_t* a32 = reinterpret_cast<uint32_t*>(a.arr);
    __shared__ unsigned long long s_tile;
    unsigned long long my_count = 0, my_parents = 0, my_children = 0, my_scan = 0;
    const uint64_t m55 = 0x5555555555555555ull;
    const uint64_t pat = (CUR == 2u) ? 0xAAAAAAAAAAAAAAAAull : 0xFFFFFFFFFFFFFFFFull;
    const uint64_t flip = CUR ^ OLD;  // 3 (CUR=2) or 2 (CUR=3)
    for (;;) {
        if (tid == 0) s_tile = atomicAdd(&st->tile_counter, 1ull);
        __syncthreads();
        const uint64_t tile = s_tile;
        __syncthreads();
        if (tile >= a.ntiles) break;
        const uint64_t w0 = tile * a.tile_words;
        const uint64_t w1 = (w0 + a.tile_words < a.nwords) ? w0 + a.tile_words : a.nwords;
        auto mark = [&](uint64_t child) {
            if (child >= a.nbits) {
                atomicOr(&st->error, 1u);
                return;
            }
            ++my_children;
            uint32_t* p = a32 + (child >> 4);
            const uint32_t sh = (uint32_t)(child & 15) * 2;
            uint32_t old = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            while (((old >> sh) & 3u) == 0u) {
                const uint32_t prev = atomicCAS(p, old, old | (NEXT << sh));
                if (prev == old) {
                    ++my_count;
                    break;
                }
                old = prev;
            }
        };
        auto run_batch = [&](uint32_t n) {
            for (uint32_t j = lane; j < n; j += 64) {
                const uint64_t parent = w0 * 32 + q[j];
                ++my_parents;
                PZ::expand(parent, 0u, a.prm, [&](int, uint64_t child, uint32_t) { mark(child); });
            }
        };
        int qn = 0;
        for (uint64_t g = w0 + (uint64_t)wave * 64; g < w1; g += (uint64_t)NW * 64) {
            const uint64_t wi = g + lane;
            uint64_t m = 0;
            if (wi < w1) {
                const uint64_t v = __hip_atomic_load(&a.arr[wi], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const uint64_t x = v ^ pat;               // CUR fields become 00
                m = ~(x | (x >> 1)) & m55;                // low bit of each CUR field
                if (m) {
                    const uint64_t fx = (flip & 1 ? m : 0) | (flip & 2 ? (m << 1) : 0);
                    uint32_t* p = reinterpret_cast<uint32_t*>(&a.arr[wi]);
                    if ((uint32_t)fx) atomicXor(p, (uint32_t)fx);
                    if ((uint32_t)(fx >> 32)) atomicXor(p + 1, (uint32_t)(fx >> 32));
                }
            }
            const uint32_t c = (uint32_t)__popcll(m);
            const uint32_t incl = wave_incl_scan(c, lane);
            const int total = (int)__shfl(incl, 63, 64);
            int off = qn + (int)(incl - c);
            int end_all = qn + total;
            const uint32_t base = (uint32_t)((wi - w0) * 32);
            for (;;) {
                while (m && off < QCAP) {
                    q[off++] = base + (uint32_t)((__ffsll((unsigned long long)m) - 1) >> 1);
                    m &= m - 1;
                }
                if (end_all < QCAP) {
                    qn = end_all;
                    break;
                }
                wave_sync();
                run_batch(QCAP);
                wave_sync();
                off -= QCAP;
                end_all -= QCAP;
                qn = 0;
                if (end_all == 0) break;
            }
        }
        wave_sync();
        run_batch((uint32_t)qn);
        __syncthreads();
    }
    block_add3<TPB>(my_count, my_parents, my_children, &st->acc_count, &st->expanded, &st->children);
}

// Whole-array hash for the two-bit search (run after a layer when hashes are wanted).
template <int TPB>
__global__ __launch_bounds__(TPB) void k_bfs2_hash(const uint64_t* arr, uint64_t nwords, BfsDevState* st) {
    if (st->done) return;
    unsigned long long h = 0;
    for (uint64_t j = (uint64_t)blockIdx.x * TPB + threadIdx.x; j < nwords; j += (uint64_t)gridDim.x * TPB)
        h += word_hash(j, arr[j]);
    block_add3<TPB>(h, 0ull, 0ull, &st->acc_hash, nullptr, nullptr);
}

// ---------------------------------------------------------------- multi-search
// Several searches (jobs) of different puzzles advance one layer per step in
// the same three launches: every block takes (job, tile) pairs from one global
// counter over the concatenated tile lists, and (job, chunk) pairs for the
// finalize; one block per job does the bookkeeping.

constexpr uint32_t MULTI_TL_WORDS = 256;  // segments tracked in LDS per job (8192)

struct MultiJob {
    ExpandArgs e;
    FinalizeArgs f;
    LayerEndArgs l;
    uint32_t kind;  // index into the multi-search puzzle table
    uint32_t slot;  // the job's slot in the done-flag array
    uint64_t tile_base;
    uint64_t chunk_base;
};

struct MultiArgs {
    const MultiJob* jobs;
    uint32_t njobs;
    uint64_t total_tiles;
    uint64_t total_chunks;
    unsigned long long* counters;  // [0] tiles taken, [1] chunks taken
    uint32_t* done;                // per slot: the job finished
};

// the job owning item g (first job whose base is <= g, bases ascending)
template <bool TILES>
__device__ __forceinline__ uint32_t multi_find(const MultiArgs& m, uint64_t g) {
    uint32_t lo = 0, hi = m.njobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        const uint64_t base = TILES ? m.jobs[mid].tile_base : m.jobs[mid].chunk_base;
        if (base <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <class KINDS, int TPB>
__global__ __launch_bounds__(TPB) void k_multi_expand(MultiArgs m, uint32_t win_u32_max) {
    extern __shared__ uint32_t smem[];
    constexpr int NW = TPB / 64;
    ExpandLds lds;
    lds.win = smem;
    lds.cwin = nullptr;  // multi-search layers run top-down
    lds.queues = smem + win_u32_max;
    lds.tl = smem + win_u32_max + NW * QCAP;
    __shared__ unsigned long long s_tile;
    __shared__ uint32_t s_local;
    for (;;) {
        if (threadIdx.x == 0) s_tile = atomicAdd(&m.counters[0], 1ull);
        __syncthreads();
        const uint64_t g = s_tile;
        if (g >= m.total_tiles) break;
        const MultiJob& J = m.jobs[multi_find<true>(m, g)];
        if (J.e.st->done) {
            __syncthreads();
            continue;
        }
        const uint32_t nseg = J.e.nseg;
        lds.tl_words = (nseg > 1 && (nseg + 31) / 32 <= MULTI_TL_WORDS) ? (nseg + 31) / 32 : 0u;
        for (uint32_t i = threadIdx.x; i < lds.tl_words; i += TPB) lds.tl[i] = 0;
        ExpandTally tally;
        KINDS::template run<TPB>(J.kind, J.e, g - J.tile_base, lds, &s_local, tally);
        for (uint32_t i = threadIdx.x; i < lds.tl_words; i += TPB) {
            const uint32_t v = lds.tl[i];
            if (v) gor(&J.e.touched[i], v);
        }
        block_add3<TPB>(tally.children, tally.parents, tally.scan, &J.e.st->children, &J.e.st->expanded,
                        &J.e.st->scan_words);
    }
}

template <int TPB>
__global__ __launch_bounds__(TPB) void k_multi_finalize(MultiArgs m) {
    __shared__ unsigned long long s_chunk;
    for (;;) {
        if (threadIdx.x == 0) s_chunk = atomicAdd(&m.counters[1], 1ull);
        __syncthreads();
        const uint64_t g = s_chunk;
        __syncthreads();
        if (g >= m.total_chunks) break;
        const MultiJob& J = m.jobs[multi_find<false>(m, g)];
        if (J.f.st->done) continue;
        FinTally tally;
        finalize_chunk<TPB>(J.f, g - J.chunk_base, tally);
        flush_fin_tally(J.f.st, tally);
    }
}

template <int TPB>
__global__ __launch_bounds__(TPB) void k_multi_layer_end(MultiArgs m) {
    const MultiJob& J = m.jobs[blockIdx.x];
    layer_end_body<TPB>(J.l);
    __syncthreads();
    if (threadIdx.x == 0) {
        m.done[J.slot] = J.l.st->done;
        if (blockIdx.x == 0) {
            m.counters[0] = 0;
            m.counters[1] = 0;
        }
    }
}

}  // namespace psk
