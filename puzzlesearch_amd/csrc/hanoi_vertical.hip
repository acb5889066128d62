// The HanoiTowers "vertical" solver on the device: the symmetry-reduced 4-peg
// Hanoi expanders of HanoiTowers/HanoiExpandGPU.cu:98-229 and the three-bit
// search of HanoiTowers/HanoiSearch.cpp:46-290 with its segments kept in HBM
// instead of files.
//
// A state index holds disk d's peg in bits 2d..2d+1.  The search keeps only
// states whose pegs 1..3 are ordered by their bottom disk, largest first and
// empty pegs last, and restores each layer's count to the unreduced puzzle
// with PegsMult (HanoiTowers/HanoiTowers.h:48-55); layer 0 counts 1.
//
// Segments hold 2^segment_bits states (2^32 in the reference).  Only segments
// that a layer reaches get device memory: three bit arrays (cur, next, old,
// rotating by layer) of 2^segment_bits bits each, 1.5 GiB for 2^32.  Per layer:
//   expand    every segment with a current frontier: each block takes tiles of
//             2^18 states, ORs the children that stay in the tile into an LDS
//             window and stores it into the tile's next words (nothing else
//             writes them in this phase); children elsewhere in the segment or
//             in another segment are appended to a list.  In a segment whose
//             disks above it hold pegs 1 and 2 (PegsMult(segment) = 6) no move
//             breaks the order of pegs 1..3, so the smallest disk's four
//             positions are set at once and only the three moves among the
//             other pegs are generated (HanoiSearch.cpp:104-110,
//             kernel_expand_inseg_without_smallest)
//   apply     the listed children are ORed into their segment's next array;
//             a segment reached for the first time is allocated and the pass
//             is repeated (ORs are idempotent)
//   finalize  new = next & ~old & ~cur per word (the moves have odd cycles),
//             counted, restored with PegsMult and hashed; old is cleared
//             (ScanBitsAndClearWithExcl, HanoiSearch.cpp:137-160)
#include <algorithm>
#include <chrono>
#include <cstring>
#include <new>
#include <vector>

#include "devbuf.h"
#include "ps_internal.h"

namespace {

constexpr int HV_MAX_DISKS = 24;
constexpr uint64_t HV_BIT_MASK = 0x5555555555555555ull;
constexpr int HV_TILE_WORDS = 4096;  // 2^18 states: moves of disks 0..8 stay in the tile
constexpr int HV_TPB = 256;
constexpr uint32_t HV_NO_SLOT = 0xFFFFFFFFu;

struct HvState {
    uint64_t pegs[4];
    int top[4];
};

__device__ __forceinline__ int hv_low(uint64_t x) { return x ? __builtin_ctzll(x) : 255; }
__device__ __forceinline__ int hv_high(uint64_t x) { return x ? 63 - __builtin_clzll(x) : -1; }

__device__ __forceinline__ void hv_from_index(uint64_t index, uint64_t mask, HvState& s) {
    const uint64_t p0 = index >> 1, p1 = index;
    s.pegs[0] = (~p0 & ~p1) & mask;
    s.pegs[1] = (~p0 & p1) & mask;
    s.pegs[2] = (p0 & ~p1) & mask;
    s.pegs[3] = (p0 & p1) & mask;
#pragma unroll
    for (int p = 0; p < 4; ++p) s.top[p] = hv_low(s.pegs[p]);
}

// the smaller top disk of pegs a, b changes peg; pegs 1..3 are put back in
// bottom-disk order when one of the two pegs was or became empty
__device__ __forceinline__ uint64_t hv_child(const HvState& s, int a, int b) {
    uint64_t pg[4] = {s.pegs[0], s.pegs[1], s.pegs[2], s.pegs[3]};
    const int disk = min(s.top[a], s.top[b]);
    const bool was_empty = pg[a] == 0 || pg[b] == 0;
    if (disk != 255) {
        const uint64_t bit = 1ull << disk;
        pg[a] ^= bit;
        pg[b] ^= bit;
    }
    if (was_empty || pg[a] == 0 || pg[b] == 0) {
        int bt[4] = {hv_high(pg[0]), hv_high(pg[1]), hv_high(pg[2]), hv_high(pg[3])};
        auto order = [&](int i, int j) {
            if (bt[i] < bt[j]) {
                const int t = bt[i];
                bt[i] = bt[j];
                bt[j] = t;
                const uint64_t u = pg[i];
                pg[i] = pg[j];
                pg[j] = u;
            }
        };
        order(2, 3);
        order(1, 2);
        order(2, 3);
    }
    return pg[1] | (pg[2] << 1) | (pg[3] * 3);
}

__constant__ int8_t c_hv_pairs[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
__constant__ int8_t c_hv_other[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};

__device__ __forceinline__ int hv_pegs_mult(uint64_t index) {
    const uint64_t p0 = index >> 1, p1 = index;
    const uint64_t pegs1 = (~p0 & p1) & HV_BIT_MASK;
    const uint64_t pegs2 = (p0 & ~p1) & HV_BIT_MASK;
    return 6 - (pegs2 == 0) * 3 - (pegs1 == 0);
}

__device__ __forceinline__ uint64_t hv_mix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// kernel_expand_inseg / kernel_expand_inseg_without_smallest: 6 (or 3) slots
// per parent, the parent's own id where the child leaves the 2^32 segment
__global__ void k_hv_expand_slots(uint32_t segment, uint64_t mask, const uint32_t* __restrict__ idx, uint64_t count,
                                  uint32_t* __restrict__ out, int without_smallest) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint32_t lo = idx[i];
    const uint64_t index = ((uint64_t)segment << 32) | lo;
    HvState s;
    hv_from_index(index, mask, s);
    if (!without_smallest) {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const uint64_t c = hv_child(s, c_hv_pairs[k][0], c_hv_pairs[k][1]);
            out[i * 6 + k] = (c >> 32) == segment ? (uint32_t)c : lo;
        }
    } else {
        const int p0 = (int)(index & 3);
        const int o0 = c_hv_other[p0][0], o1 = c_hv_other[p0][1], o2 = c_hv_other[p0][2];
        const uint64_t c0 = hv_child(s, o0, o1), c1 = hv_child(s, o0, o2), c2 = hv_child(s, o1, o2);
        out[i * 3 + 0] = (c0 >> 32) == segment ? (uint32_t)c0 : lo;
        out[i * 3 + 1] = (c1 >> 32) == segment ? (uint32_t)c1 : lo;
        out[i * 3 + 2] = (c2 >> 32) == segment ? (uint32_t)c2 : lo;
    }
}

// HanoiTowers::Expand(index) (HanoiTowers/HanoiTowers.cpp:152-172): the six
// children of each 64-bit index, moves (0,1) (0,2) (0,3) (1,2) (1,3) (2,3)
__global__ void k_hv_expand_full(uint64_t mask, const uint64_t* __restrict__ idx, uint64_t count,
                                 uint64_t* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    HvState s;
    hv_from_index(idx[i], mask, s);
#pragma unroll
    for (int k = 0; k < 6; ++k) out[i * 6 + k] = hv_child(s, c_hv_pairs[k][0], c_hv_pairs[k][1]);
}

struct HvExpandArgs {
    const unsigned long long* cur;
    unsigned long long* next;
    uint64_t nwords;      // words of the segment
    uint64_t tile_words;  // words per tile (a power of two, at most HV_TILE_WORDS)
    uint64_t ntiles;
    uint64_t segment;
    int seg_bits;
    int full_sym;
    uint64_t mask;                 // the disks' bits of an index (BIT_MASK restricted)
    unsigned long long* list;      // children outside their parent's tile (global indexes)
    uint64_t list_cap;
    unsigned long long* counters;  // [0] listed children, [1] parents expanded, [2] children
};

__device__ __forceinline__ void hv_emit(const HvExpandArgs& a, uint32_t* win, uint64_t tile_base_bit, uint64_t tile_bits,
                                        uint64_t child, unsigned& nlist, unsigned& nchild) {
    ++nchild;
    if ((child >> a.seg_bits) == a.segment) {
        const uint64_t local = child & ((1ull << a.seg_bits) - 1);
        const uint64_t off = local - tile_base_bit;
        if (off < tile_bits) {
            atomicOr(&win[off >> 5], 1u << (off & 31));
            return;
        }
    }
    const unsigned long long slot = atomicAdd(&a.counters[0], 1ull);
    if (slot < a.list_cap) a.list[slot] = child;
    ++nlist;
}

__global__ __launch_bounds__(HV_TPB) void k_hv_expand(HvExpandArgs a) {
    extern __shared__ uint32_t win[];  // tile_words * 2 u32
    __shared__ unsigned long long s_par, s_child;
    if (threadIdx.x == 0) s_par = s_child = 0;
    const uint64_t tile_bits = a.tile_words * 64;
    const uint64_t seg_base = a.segment << a.seg_bits;
    unsigned npar = 0, nchild = 0, nlist = 0;
    for (uint64_t t = blockIdx.x; t < a.ntiles; t += gridDim.x) {
        for (uint64_t j = threadIdx.x; j < a.tile_words * 2; j += blockDim.x) win[j] = 0u;
        __syncthreads();
        const uint64_t w0 = t * a.tile_words;
        const uint64_t tile_base_bit = w0 * 64;
        for (uint64_t j = threadIdx.x; j < a.tile_words; j += blockDim.x) {
            uint64_t w = a.cur[w0 + j];
            while (w) {
                const int b = __builtin_ctzll(w);
                w &= w - 1;
                const uint64_t local = (w0 + j) * 64 + (uint64_t)b;
                const uint64_t index = seg_base | local;
                ++npar;
                HvState s;
                hv_from_index(index, a.mask, s);
                if (a.full_sym) {
                    // the smallest disk on each of the four pegs: bits local&~3 .. +3,
                    // the parent itself among them (excluded by the finalize)
                    const uint64_t off = (local & ~3ull) - tile_base_bit;
                    atomicOr(&win[off >> 5], 0xFu << (off & 31));
                    nchild += 3;
                    const int p0 = (int)(index & 3);
                    const int o0 = c_hv_other[p0][0], o1 = c_hv_other[p0][1], o2 = c_hv_other[p0][2];
                    hv_emit(a, win, tile_base_bit, tile_bits, hv_child(s, o0, o1), nlist, nchild);
                    hv_emit(a, win, tile_base_bit, tile_bits, hv_child(s, o0, o2), nlist, nchild);
                    hv_emit(a, win, tile_base_bit, tile_bits, hv_child(s, o1, o2), nlist, nchild);
                } else {
#pragma unroll
                    for (int k = 0; k < 6; ++k)
                        hv_emit(a, win, tile_base_bit, tile_bits, hv_child(s, c_hv_pairs[k][0], c_hv_pairs[k][1]),
                                nlist, nchild);
                }
            }
        }
        __syncthreads();
        // the tile's next words: only this block writes them while the expand runs
        for (uint64_t j = threadIdx.x; j < a.tile_words; j += blockDim.x) {
            const uint64_t v = (uint64_t)win[2 * j] | ((uint64_t)win[2 * j + 1] << 32);
            if (v) a.next[w0 + j] = v;
        }
        __syncthreads();
    }
    (void)nlist;
    atomicAdd(&s_par, (unsigned long long)npar);
    atomicAdd(&s_child, (unsigned long long)nchild);
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(&a.counters[1], s_par);
        atomicAdd(&a.counters[2], s_child);
    }
}

struct HvApplyArgs {
    const unsigned long long* list;
    uint64_t count;
    const uint32_t* slot_of_seg;      // [nseg]
    unsigned long long* const* next;  // [slot] this layer's next array
    int seg_bits;
    uint64_t nseg;
    uint8_t* need;                    // [nseg] set where a child's segment has no memory yet
    unsigned long long* missing;
};

__global__ __launch_bounds__(HV_TPB) void k_hv_apply(HvApplyArgs a) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.count; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t c = a.list[i];
        const uint64_t seg = c >> a.seg_bits;
        if (seg >= a.nseg) continue;  // cannot happen for states of the puzzle
        const uint32_t slot = a.slot_of_seg[seg];
        if (slot == HV_NO_SLOT) {
            a.need[seg] = 1;
            atomicAdd(a.missing, 1ull);
            continue;
        }
        const uint64_t local = c & ((1ull << a.seg_bits) - 1);
        atomicOr(&a.next[slot][local >> 6], 1ull << (local & 63));
    }
}

struct HvFinalizeArgs {
    unsigned long long* next;
    unsigned long long* old;
    const unsigned long long* cur;
    uint64_t nwords;
    uint64_t segment;
    int seg_bits;
    int full_sym;
    unsigned long long* out;  // [0] count, [1] restored, [2] hash
};

__global__ __launch_bounds__(HV_TPB) void k_hv_finalize(HvFinalizeArgs a) {
    __shared__ unsigned long long s_sum[3];
    if (threadIdx.x < 3) s_sum[threadIdx.x] = 0;
    __syncthreads();
    unsigned long long cnt = 0, res = 0, hash = 0;
    const uint64_t seg_base = a.segment << a.seg_bits;
    const uint64_t gw0 = seg_base >> 6;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < a.nwords; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t nx = a.next[j], od = a.old[j], cu = a.cur[j];
        const uint64_t n = nx & ~od & ~cu;
        if (nx != n) a.next[j] = n;
        if (od) a.old[j] = 0;
        if (!n) continue;
        const int pc = __popcll(n);
        cnt += (unsigned long long)pc;
        hash += hv_mix(hv_mix(gw0 + j) ^ n);
        if (a.full_sym) {
            res += 6ull * (unsigned long long)pc;
        } else {
            uint64_t w = n;
            while (w) {
                const int b = __builtin_ctzll(w);
                w &= w - 1;
                res += (unsigned long long)hv_pegs_mult(seg_base | (j * 64 + (uint64_t)b));
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        cnt += __shfl_down(cnt, off);
        res += __shfl_down(res, off);
        hash += __shfl_down(hash, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&s_sum[0], cnt);
        atomicAdd(&s_sum[1], res);
        atomicAdd(&s_sum[2], hash);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(&a.out[0], s_sum[0]);
        atomicAdd(&a.out[1], s_sum[1]);
        atomicAdd(&a.out[2], s_sum[2]);
    }
}

__global__ void k_hv_set_bit(unsigned long long* words, uint64_t bit) {
    if (threadIdx.x == 0 && blockIdx.x == 0) words[bit >> 6] |= 1ull << (bit & 63);
}

uint64_t host_mix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

int host_pegs_mult(uint64_t index) {
    const uint64_t p0 = index >> 1, p1 = index;
    const uint64_t pegs1 = (~p0 & p1) & HV_BIT_MASK;
    const uint64_t pegs2 = (p0 & ~p1) & HV_BIT_MASK;
    return 6 - (pegs2 == 0) * 3 - (pegs1 == 0);
}

// pegs 1..3 ordered by bottom disk, largest first, empty pegs last
bool host_reduced(int disks, uint64_t index) {
    int bottom[4] = {-1, -1, -1, -1};
    for (int d = 0; d < disks; ++d) bottom[(index >> (2 * d)) & 3] = d;
    return bottom[1] >= bottom[2] && bottom[2] >= bottom[3];
}

uint64_t disks_mask(int disks) { return HV_BIT_MASK & ((1ull << (2 * disks)) - 1); }

}  // namespace

struct ps_hanoi_search {
    int disks = 0, seg_bits = 0;
    uint64_t nseg = 0, nwords = 0, tile_words = 0, ntiles = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    uint64_t budget = 0, used = 0;
    struct Slot {
        uint64_t segment;
        unsigned long long* arr[3];
        uint64_t cur_count;
    };
    std::vector<Slot> slots;
    std::vector<uint32_t> slot_of_seg;  // host copy
    uint32_t* d_slot_of_seg = nullptr;
    unsigned long long** d_next = nullptr;    // [slots capacity]
    unsigned long long** h_next = nullptr;    // pinned
    uint32_t next_cap = 0;
    uint8_t* d_need = nullptr;
    unsigned long long* d_counters = nullptr;  // [0..2] expand, [3] missing, then 3 per slot
    unsigned long long* h_counters = nullptr;  // pinned
    uint64_t counters_cap = 0;                 // u64 entries
    unsigned long long* d_list = nullptr;
    uint64_t list_cap = 0;
    uint32_t grid_expand = 0, grid_fin = 0;
    ps_hanoi_search_stats stats{};
};

namespace ps {
namespace {

int destroy_hv(ps_hanoi_search* h) {
    if (!h) return PS_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (auto& s : h->slots)
        for (auto* p : s.arr)
            if (p) (void)hipFree(p);
    if (h->d_slot_of_seg) (void)hipFree(h->d_slot_of_seg);
    if (h->d_next) (void)hipFree(h->d_next);
    if (h->h_next) (void)hipHostFree(h->h_next);
    if (h->d_need) (void)hipFree(h->d_need);
    if (h->d_counters) (void)hipFree(h->d_counters);
    if (h->h_counters) (void)hipHostFree(h->h_counters);
    if (h->d_list) (void)hipFree(h->d_list);
    if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    (void)hipGetLastError();
    return PS_OK;
}

// room for per-slot counters and next pointers of `n` slots
int ensure_slot_tables(ps_hanoi_search* h, uint32_t n) {
    if (n > h->next_cap) {
        uint32_t cap = std::max<uint32_t>(16, h->next_cap);
        while (cap < n) cap *= 2;
        if (h->d_next) (void)hipFree(h->d_next);
        if (h->h_next) (void)hipHostFree(h->h_next);
        h->d_next = nullptr;
        h->h_next = nullptr;
        PS_HIP_CHECK(hipMalloc((void**)&h->d_next, cap * sizeof(void*)));
        PS_HIP_CHECK(hipHostMalloc((void**)&h->h_next, cap * sizeof(void*), hipHostMallocDefault));
        h->next_cap = cap;
    }
    const uint64_t want = 4 + 3ull * n;
    if (want > h->counters_cap) {
        uint64_t cap = std::max<uint64_t>(64, h->counters_cap);
        while (cap < want) cap *= 2;
        if (h->d_counters) (void)hipFree(h->d_counters);
        if (h->h_counters) (void)hipHostFree(h->h_counters);
        h->d_counters = nullptr;
        h->h_counters = nullptr;
        PS_HIP_CHECK(hipMalloc((void**)&h->d_counters, cap * 8));
        PS_HIP_CHECK(hipHostMalloc((void**)&h->h_counters, cap * 8, hipHostMallocDefault));
        h->counters_cap = cap;
    }
    return PS_OK;
}

int add_slot(ps_hanoi_search* h, uint64_t segment) {
    const uint64_t bytes = h->nwords * 8;
    if (h->used + 3 * bytes > h->budget)
        return fail(PS_ERR_CAPACITY, "segment %llu needs %llu more bytes of HBM; the budget is %llu (%zu segments held)",
                    (unsigned long long)segment, (unsigned long long)(3 * bytes), (unsigned long long)h->budget,
                    h->slots.size());
    ps_hanoi_search::Slot s{};
    s.segment = segment;
    for (int r = 0; r < 3; ++r) {
        if (hipMalloc((void**)&s.arr[r], bytes) != hipSuccess) {
            (void)hipGetLastError();
            for (int q = 0; q < r; ++q) (void)hipFree(s.arr[q]);
            return fail(PS_ERR_OUT_OF_MEMORY, "device allocation of a %llu-byte segment array failed",
                        (unsigned long long)bytes);
        }
        if (int rc = launch_fill(s.arr[r], 0, bytes, h->stream)) return rc;
    }
    h->used += 3 * bytes;
    h->slot_of_seg[segment] = (uint32_t)h->slots.size();
    h->slots.push_back(s);
    if (int rc = ensure_slot_tables(h, (uint32_t)h->slots.size())) return rc;
    h->stats.segments_allocated = h->slots.size();
    return PS_OK;
}

int upload_tables(ps_hanoi_search* h, int layer) {
    for (size_t i = 0; i < h->slots.size(); ++i) h->h_next[i] = h->slots[i].arr[(layer + 1) % 3];
    PS_HIP_CHECK(hipMemcpyAsync(h->d_next, h->h_next, h->slots.size() * sizeof(void*), hipMemcpyHostToDevice,
                                h->stream));
    PS_HIP_CHECK(hipMemcpyAsync(h->d_slot_of_seg, h->slot_of_seg.data(), h->nseg * 4, hipMemcpyHostToDevice,
                                h->stream));
    // the pinned pointer table is read by the copy: wait before it changes
    PS_HIP_CHECK(hipStreamSynchronize(h->stream));
    return PS_OK;
}

int read_counters(ps_hanoi_search* h, uint64_t n) {
    PS_HIP_CHECK(hipMemcpyAsync(h->h_counters, h->d_counters, n * 8, hipMemcpyDeviceToHost, h->stream));
    PS_HIP_CHECK(hipStreamSynchronize(h->stream));
    return PS_OK;
}

int expand_all(ps_hanoi_search* h, int layer, uint64_t* expanded_segments) {
    if (int r = launch_fill(h->d_counters, 0, 4 * 8, h->stream)) return r;
    *expanded_segments = 0;
    for (auto& s : h->slots) {
        if (!s.cur_count) continue;
        ++*expanded_segments;
        HvExpandArgs a{};
        a.cur = s.arr[layer % 3];
        a.next = s.arr[(layer + 1) % 3];
        a.nwords = h->nwords;
        a.tile_words = h->tile_words;
        a.ntiles = h->ntiles;
        a.segment = s.segment;
        a.seg_bits = h->seg_bits;
        a.full_sym = host_pegs_mult(s.segment) == 6 && h->seg_bits < 2 * h->disks;
        a.mask = disks_mask(h->disks);
        a.list = h->d_list;
        a.list_cap = h->list_cap;
        a.counters = h->d_counters;
        const uint32_t grid = (uint32_t)std::min<uint64_t>(h->grid_expand, h->ntiles);
        hipLaunchKernelGGL(k_hv_expand, dim3(grid), dim3(HV_TPB), h->tile_words * 8, h->stream, a);
        PS_HIP_CHECK(hipGetLastError());
    }
    return PS_OK;
}

int run_layer(ps_hanoi_search* h, int layer, uint64_t* count, uint64_t* restored, uint64_t* hash) {
    // expand; repeated with a longer list if it overflowed (the window stores
    // and list entries are rewritten identically)
    uint64_t xsegs = 0;
    if (int r = upload_tables(h, layer)) return r;
    if (int r = expand_all(h, layer, &xsegs)) return r;
    if (int r = read_counters(h, 3)) return r;
    uint64_t listed = h->h_counters[0];
    if (listed > h->list_cap) {
        uint64_t cap = std::max<uint64_t>(h->list_cap, 1 << 16);
        while (cap < listed) cap *= 2;
        if (h->d_list) (void)hipFree(h->d_list);
        h->d_list = nullptr;
        h->list_cap = 0;
        PS_HIP_CHECK(hipMalloc((void**)&h->d_list, cap * 8));
        h->list_cap = cap;
        if (int r = expand_all(h, layer, &xsegs)) return r;
        if (int r = read_counters(h, 3)) return r;
        listed = h->h_counters[0];
        if (listed > h->list_cap) return fail(PS_ERR_INTERNAL, "child list grew between two identical expansions");
    }
    const uint64_t layer_children = h->h_counters[2];
    h->stats.expanded += h->h_counters[1];
    h->stats.children += layer_children;
    h->stats.listed_children += listed;
    // apply the listed children; allocate the segments they reach first
    for (int round = 0; listed > 0; ++round) {
        if (int r = launch_fill(h->d_counters + 3, 0, 8, h->stream)) return r;
        HvApplyArgs ap{};
        ap.list = h->d_list;
        ap.count = listed;
        ap.slot_of_seg = h->d_slot_of_seg;
        ap.next = h->d_next;
        ap.seg_bits = h->seg_bits;
        ap.nseg = h->nseg;
        ap.need = h->d_need;
        ap.missing = h->d_counters + 3;
        const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((listed + HV_TPB - 1) / HV_TPB,
                                                                                 (uint64_t)hip_cu_count() * 8));
        hipLaunchKernelGGL(k_hv_apply, dim3(grid), dim3(HV_TPB), 0, h->stream, ap);
        PS_HIP_CHECK(hipGetLastError());
        PS_HIP_CHECK(hipMemcpyAsync(h->h_counters + 3, h->d_counters + 3, 8, hipMemcpyDeviceToHost, h->stream));
        PS_HIP_CHECK(hipStreamSynchronize(h->stream));
        if (h->h_counters[3] == 0) break;
        if (round > 0) return fail(PS_ERR_INTERNAL, "segments still missing after their allocation");
        std::vector<uint8_t> need(h->nseg);
        PS_HIP_CHECK(hipMemcpy(need.data(), h->d_need, h->nseg, hipMemcpyDeviceToHost));
        for (uint64_t sgm = 0; sgm < h->nseg; ++sgm)
            if (need[sgm] && h->slot_of_seg[sgm] == HV_NO_SLOT)
                if (int r = add_slot(h, sgm)) return r;
        if (int r = launch_fill(h->d_need, 0, h->nseg, h->stream)) return r;
        if (int r = upload_tables(h, layer)) return r;
    }
    // finalize every segment held
    const uint64_t nslots = h->slots.size();
    if (int r = launch_fill(h->d_counters + 4, 0, 3 * nslots * 8, h->stream)) return r;
    for (uint64_t i = 0; i < nslots; ++i) {
        auto& s = h->slots[i];
        HvFinalizeArgs f{};
        f.next = s.arr[(layer + 1) % 3];
        f.old = s.arr[(layer + 2) % 3];
        f.cur = s.arr[layer % 3];
        f.nwords = h->nwords;
        f.segment = s.segment;
        f.seg_bits = h->seg_bits;
        f.full_sym = host_pegs_mult(s.segment) == 6 && h->seg_bits < 2 * h->disks;
        f.out = h->d_counters + 4 + 3 * i;
        hipLaunchKernelGGL(k_hv_finalize, dim3(h->grid_fin), dim3(HV_TPB), 0, h->stream, f);
        PS_HIP_CHECK(hipGetLastError());
    }
    if (int r = read_counters(h, 4 + 3 * nslots)) return r;
    uint64_t c = 0, rs = 0, hs = 0;
    for (uint64_t i = 0; i < nslots; ++i) {
        h->slots[i].cur_count = h->h_counters[4 + 3 * i];
        c += h->h_counters[4 + 3 * i];
        rs += h->h_counters[5 + 3 * i];
        hs += h->h_counters[6 + 3 * i];
    }
    // algorithmic bytes: 8 B per cur word of an expanded segment and per child
    // marked, 16 B more per listed child (written, read back), 40 B per
    // finalized word (next, old, cur read; next, old written)
    h->stats.algorithmic_bytes += 8ull * h->nwords * xsegs + 8ull * layer_children + 16ull * listed +
                                  40ull * h->nwords * nslots;
    *count = c;
    *restored = rs;
    *hash = hs;
    return PS_OK;
}

}  // namespace
}  // namespace ps

using namespace ps;

extern "C" {

int ps_hanoi_expand_in_segment(int32_t disks, uint32_t segment, const uint32_t* d_idx, uint64_t count,
                               uint32_t* d_children, int32_t without_smallest, ps_stream s) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (disks < 1 || disks > HV_MAX_DISKS) return fail(PS_ERR_INVALID_ARGUMENT, "disks must be in 1..%d", HV_MAX_DISKS);
    if (count == 0) return PS_OK;
    if (!d_idx || !d_children) return fail(PS_ERR_INVALID_ARGUMENT, "NULL buffer");
    if (2 * disks <= 32 ? segment != 0 : ((uint64_t)segment >> (2 * disks - 32)) != 0)
        return fail(PS_ERR_INDEX_RANGE, "segment %u outside the %d-disk index space", segment, disks);
    const uint64_t blocks = (count + 255) / 256;
    if (blocks > 0x7FFFFFFFull) return fail(PS_ERR_INVALID_ARGUMENT, "count too large");
    hipLaunchKernelGGL(k_hv_expand_slots, dim3((uint32_t)blocks), dim3(256), 0, stream_of(s), segment,
                       disks_mask(disks), d_idx, count, d_children, without_smallest ? 1 : 0);
    PS_HIP_CHECK(hipGetLastError());
    return PS_OK;
    PS_API_END
}

int ps_hanoi_expand(int32_t disks, const uint64_t* d_idx, uint64_t count, uint64_t* d_children, ps_stream s) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (disks < 1 || disks > HV_MAX_DISKS) return fail(PS_ERR_INVALID_ARGUMENT, "disks must be in 1..%d", HV_MAX_DISKS);
    if (count == 0) return PS_OK;
    if (!d_idx || !d_children) return fail(PS_ERR_INVALID_ARGUMENT, "NULL buffer");
    const uint64_t blocks = (count + 255) / 256;
    if (blocks > 0x7FFFFFFFull) return fail(PS_ERR_INVALID_ARGUMENT, "count too large");
    hipLaunchKernelGGL(k_hv_expand_full, dim3((uint32_t)blocks), dim3(256), 0, stream_of(s), disks_mask(disks), d_idx,
                       count, d_children);
    PS_HIP_CHECK(hipGetLastError());
    return PS_OK;
    PS_API_END
}

int ps_hanoi_search_create(int32_t disks, int32_t segment_bits, uint64_t hbm_budget, ps_stream s,
                           ps_hanoi_search** out) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (!out) return fail(PS_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    if (disks < 3 || disks > HV_MAX_DISKS) return fail(PS_ERR_INVALID_ARGUMENT, "disks must be in 3..%d", HV_MAX_DISKS);
    if (segment_bits == 0) segment_bits = 32;
    if (segment_bits < 8 || segment_bits > 32 || (segment_bits & 1))
        return fail(PS_ERR_INVALID_ARGUMENT, "segment_bits must be even and in 8..32");
    auto* h = new (std::nothrow) ps_hanoi_search();
    if (!h) return fail(PS_ERR_OUT_OF_MEMORY, "host allocation failed");
    auto cleanup = [&](int rc) {
        destroy_hv(h);
        return rc;
    };
    h->disks = disks;
    h->seg_bits = std::min(segment_bits, 2 * disks);
    h->nseg = 1ull << (2 * disks - h->seg_bits);
    if (h->nseg > (1ull << 24)) return cleanup(fail(PS_ERR_INVALID_ARGUMENT, "more than 2^24 segments"));
    h->nwords = (1ull << h->seg_bits) / 64;
    h->tile_words = std::min<uint64_t>(HV_TILE_WORDS, h->nwords);
    h->ntiles = h->nwords / h->tile_words;
    if (s) {
        h->stream = stream_of(s);
    } else {
        if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
            return cleanup(fail(PS_ERR_HIP, "stream creation failed"));
        h->own_stream = true;
    }
    if (hbm_budget) {
        h->budget = hbm_budget;
    } else {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return cleanup(fail(PS_ERR_HIP, "hipMemGetInfo failed"));
        h->budget = free_b / 4 * 3;
    }
    h->slot_of_seg.assign(h->nseg, HV_NO_SLOT);
    if (hipMalloc((void**)&h->d_slot_of_seg, h->nseg * 4) != hipSuccess ||
        hipMalloc((void**)&h->d_need, h->nseg) != hipSuccess) {
        (void)hipGetLastError();
        return cleanup(fail(PS_ERR_OUT_OF_MEMORY, "device allocation failed"));
    }
    if (int r = launch_fill(h->d_need, 0, h->nseg, h->stream)) return cleanup(r);
    if (int r = ensure_slot_tables(h, 1)) return cleanup(r);
    h->list_cap = 1 << 20;
    if (hipMalloc((void**)&h->d_list, h->list_cap * 8) != hipSuccess) {
        (void)hipGetLastError();
        return cleanup(fail(PS_ERR_OUT_OF_MEMORY, "device allocation failed"));
    }
    const int cus = hip_cu_count();
    h->grid_expand = (uint32_t)cus * 4;
    h->grid_fin = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((h->nwords + HV_TPB - 1) / HV_TPB, (uint64_t)cus * 8));
    *out = h;
    return PS_OK;
    PS_API_END
}

int ps_hanoi_search_destroy(ps_hanoi_search* h) {
    PS_API_BEGIN
    return destroy_hv(h);
    PS_API_END
}

int ps_hanoi_search_run(ps_hanoi_search* h, uint64_t start, int32_t max_layers, uint64_t* restored,
                        uint64_t* reduced, uint64_t* hashes, int32_t cap, int32_t* n_layers) {
    PS_API_BEGIN
    if (!h || !restored || !reduced || !n_layers) return fail(PS_ERR_INVALID_ARGUMENT, "NULL argument");
    *n_layers = 0;
    if (max_layers < 1 || cap < 1) return fail(PS_ERR_INVALID_ARGUMENT, "max_layers and cap must be positive");
    if ((start >> (2 * h->disks)) != 0)
        return fail(PS_ERR_INDEX_RANGE, "start %llu outside the %d-disk index space", (unsigned long long)start,
                    h->disks);
    if (!host_reduced(h->disks, start))
        return fail(PS_ERR_INVALID_ARGUMENT, "start %llu is not a reduced state (pegs 1..3 by bottom disk, largest first)",
                    (unsigned long long)start);
    const auto t0 = std::chrono::steady_clock::now();
    // every segment held starts empty
    for (auto& s : h->slots) {
        for (auto* p : s.arr)
            if (int r = launch_fill(p, 0, h->nwords * 8, h->stream)) return r;
        s.cur_count = 0;
    }
    h->stats = ps_hanoi_search_stats{};
    h->stats.segments_allocated = h->slots.size();
    const uint64_t seg0 = start >> h->seg_bits;
    if (h->slot_of_seg[seg0] == HV_NO_SLOT)
        if (int r = add_slot(h, seg0)) return r;
    auto& s0 = h->slots[h->slot_of_seg[seg0]];
    const uint64_t local = start & ((1ull << h->seg_bits) - 1);
    hipLaunchKernelGGL(k_hv_set_bit, dim3(1), dim3(64), 0, h->stream, s0.arr[0], local);
    PS_HIP_CHECK(hipGetLastError());
    s0.cur_count = 1;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    PS_HIP_CHECK(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(PS_ERR_HIP, "event creation failed");
    }
    int rc = PS_OK;
    int layers = 1;
    restored[0] = 1;
    reduced[0] = 1;
    if (hashes) hashes[0] = host_mix(host_mix(start >> 6) ^ (1ull << (start & 63)));
    if (hipEventRecord(e0, h->stream) != hipSuccess) rc = fail(PS_ERR_HIP, "event record failed");
    while (rc == PS_OK && layers < max_layers) {
        uint64_t c = 0, r = 0, hs = 0;
        rc = run_layer(h, layers - 1, &c, &r, &hs);
        if (rc != PS_OK || c == 0) break;
        if (layers < cap) {
            restored[layers] = r;
            reduced[layers] = c;
            if (hashes) hashes[layers] = hs;
        }
        ++layers;
    }
    if (rc == PS_OK && hipEventRecord(e1, h->stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, e0, e1) == hipSuccess) h->stats.device_ms = ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != PS_OK) return rc;
    h->stats.layers = (uint64_t)layers;
    h->stats.hbm_bytes = h->used;
    h->stats.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *n_layers = std::min(layers, cap);
    return PS_OK;
    PS_API_END
}

int ps_hanoi_search_stats_get(const ps_hanoi_search* h, ps_hanoi_search_stats* st) {
    if (!h || !st) return fail(PS_ERR_INVALID_ARGUMENT, "NULL argument");
    *st = h->stats;
    return PS_OK;
}

}  // extern "C"
