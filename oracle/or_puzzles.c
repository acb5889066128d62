/*
 * or_puzzles.c -- CPU restatement of the reference puzzles' index encodings and
 * move generators (the oracle; test infrastructure only, see oracle.h).
 */
#include "oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ hashing */

uint64_t or_splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

uint64_t or_bits_hash(const uint64_t* words, uint64_t nwords) {
    uint64_t h = 0;
    for (uint64_t j = 0; j < nwords; j++) {
        if (words[j]) h += or_splitmix64(or_splitmix64(j) ^ words[j]);
    }
    return h;
}

uint64_t or_children_hash(const uint64_t* children, const int32_t* ops, uint64_t count) {
    uint64_t h = 0;
    for (uint64_t i = 0; i < count; i++) {
        if (children[i] == OR_INVALID_INDEX) continue;
        h += or_splitmix64((children[i] << 4) ^ (uint64_t)(uint32_t)ops[i]);
    }
    return h;
}

/* ------------------------------------------------- permutation rank/unrank */

/* Puzzles/PermutationMap.cpp:7-19: digit_i = tile - #(smaller tiles seen before). */
void or_perm_compact(int* arr, int size) {
    uint32_t seen = 0;
    for (int i = 0; i < size; i++) {
        int tile = arr[i];
        arr[i] = tile - __builtin_popcount(seen & ((1u << tile) - 1u));
        seen |= 1u << tile;
    }
}

/* Puzzles/PermutationMap.cpp:21-37: take the digit-th remaining tile from the
 * nibble list 0xFEDCBA9876543210 and remove it. */
void or_perm_uncompact(int* arr, int size) {
    uint64_t tiles = 0xFEDCBA9876543210ull;
    for (int i = 0; i < size; i++) {
        int d = arr[i];
        arr[i] = (int)((tiles >> (d * 4)) & 15);
        uint64_t hi = (d * 4 + 4 >= 64) ? 0 : (tiles >> (d * 4 + 4)) << (d * 4);
        uint64_t lo = tiles & ((1ull << (d * 4)) - 1ull);
        tiles = hi | lo;
    }
}

/* Puzzles/PermutationMap.cpp:39-49 */
uint64_t or_perm_rank(int* arr, int size) {
    or_perm_compact(arr, size);
    uint64_t index = 0;
    for (int i = 0; i < size; i++) index = index * (uint64_t)(size - i) + (uint64_t)arr[i];
    return index;
}

/* Puzzles/PermutationMap.cpp:51-58 */
void or_perm_unrank(uint64_t index, int* arr, int size) {
    for (int i = size - 1; i >= 0; i--) {
        arr[i] = (int)(index % (uint64_t)(size - i));
        index /= (uint64_t)(size - i);
    }
    or_perm_uncompact(arr, size);
}

/* ----------------------------------------------------------------- pancake */

/* Puzzles/PancakeOptimized.cpp:26-52 (n <= 12: plain rank; n > 12: segment =
 * rank of the first n-12 digits, low 29 bits = rank of the last 12 digits). */
uint64_t or_pancake_rank(int* arr, int n) {
    or_perm_compact(arr, n);
    if (n <= 12) {
        uint64_t index = 0;
        for (int i = 0; i < n; i++) index = index * (uint64_t)(n - i) + (uint64_t)arr[i];
        return index;
    }
    uint64_t index = 0;
    for (int i = n - 12; i < n; i++) index = index * (uint64_t)(n - i) + (uint64_t)arr[i];
    uint64_t segment = 0;
    for (int i = 0; i < n - 12; i++) segment = segment * (uint64_t)(n - i) + (uint64_t)arr[i];
    return (segment << 29) | index;
}

/* Puzzles/PancakeOptimized.cpp:54-77 */
void or_pancake_unrank(uint64_t index, int* arr, int n) {
    if (n <= 12) {
        for (int i = n - 1; i >= 0; i--) {
            arr[i] = (int)(index % (uint64_t)(n - i));
            index /= (uint64_t)(n - i);
        }
    } else {
        uint64_t segment = index >> 29;
        index &= (1ull << 29) - 1ull;
        for (int i = n - 13; i >= 0; i--) {
            arr[i] = (int)(segment % (uint64_t)(n - i));
            segment /= (uint64_t)(n - i);
        }
        for (int i = n - 1; i >= n - 12; i--) {
            arr[i] = (int)(index % (uint64_t)(n - i));
            index /= (uint64_t)(n - i);
        }
    }
    or_perm_uncompact(arr, n);
}

/* Puzzles/PancakeOptimized.cpp:12-24: reverse the last `count` elements. */
static void pancake_flip(int* arr, int size, int count) {
    int s = size - count, e = size - 1;
    while (s < e) {
        int t = arr[s];
        arr[s] = arr[e];
        arr[e] = t;
        s++;
        e--;
    }
}

/* Slots of kernel_pancake_expand / _expandInSegment / _expandCrossSegment
 * (Puzzles/PancakeGpu.cu:92-177). */
static void pancake_slots(int n, int inv_index, int mode, uint64_t packed, uint64_t* out) {
    const int max_op = n - 1;
    const int op_bits = (int)(packed & ((1ull << max_op) - 1ull));
    const uint64_t index = packed >> max_op;
    int arr[16], child[16];
    or_pancake_unrank(index, arr, n);
    int op_lo = 0, op_hi = max_op;
    if (mode == OR_EXPAND_IN_SEGMENT) { op_lo = 0; op_hi = 11; }
    if (mode == OR_EXPAND_CROSS_SEGMENT) { op_lo = 11; op_hi = max_op; }
    for (int op = op_lo; op < op_hi; op++) {
        uint64_t r = OR_INVALID_INDEX;
        if (!(op_bits & (1 << op))) {
            memcpy(child, arr, sizeof(arr));
            pancake_flip(child, n, op + 2);
            if (mode == OR_EXPAND_CROSS_SEGMENT && n > 12 && inv_index) pancake_flip(child, n, 12);
            r = (or_pancake_rank(child, n) << max_op) | (uint64_t)op;
        }
        out[op - op_lo] = r;
    }
}

/* Puzzles/PancakeGpu.cu:179-193: a child of a cross-segment op with the inverse
 * index is re-ranked with its last 12 elements reversed; a result outside the
 * segment becomes 0x1FFFFFFF (INVALID masked to 29 bits). */
void or_pancake_xseg_postprocess(int n, int segment, uint32_t* idx, uint64_t count) {
    const uint64_t seg_mask = (1ull << 29) - 1ull;
    for (uint64_t i = 0; i < count; i++) {
        uint64_t index = ((uint64_t)segment << 29) | idx[i];
        int arr[16];
        or_pancake_unrank(index, arr, n);
        pancake_flip(arr, n, 12);
        index = or_pancake_rank(arr, n);
        if ((uint64_t)segment != (index >> 29)) index = OR_INVALID_INDEX;
        idx[i] = (uint32_t)(index & seg_mask);
    }
}

/* ------------------------------------------------------ sliding tile simple */

enum { OP_UP = 0, OP_LEFT = 1, OP_RIGHT = 2, OP_DOWN = 3 };

/* kernel_sliding_tile_simple (Puzzles/SlidingPuzzleGpu.cu:25-81); the CPU twin
 * Puzzles/SlidingTilePuzzleSimple.cpp:61-99 emits the same children in the
 * same order. */
static void sliding_slots(int width, int size, uint64_t packed, uint64_t* out) {
    const int op_bits = (int)(packed & 15);
    const uint64_t index = packed >> 4;
    int arr[16], child[16];
    or_perm_unrank(index, arr, size);
    int blank = -1;
    for (int i = 0; i < size; i++)
        if (arr[i] == 0) blank = i;
    const int target[4] = {blank - width, blank + width, blank - 1, blank + 1};
    const int ok[4] = {blank >= width && !(op_bits & (1 << OP_UP)),
                       blank < size - width && !(op_bits & (1 << OP_DOWN)),
                       blank % width > 0 && !(op_bits & (1 << OP_LEFT)),
                       blank % width < width - 1 && !(op_bits & (1 << OP_RIGHT))};
    const int label[4] = {OP_DOWN, OP_UP, OP_RIGHT, OP_LEFT};
    for (int k = 0; k < 4; k++) {
        if (!ok[k]) {
            out[k] = OR_INVALID_INDEX;
            continue;
        }
        memcpy(child, arr, sizeof(arr));
        child[blank] = arr[target[k]];
        child[target[k]] = arr[blank];
        out[k] = (or_perm_rank(child, size) << 4) | (uint64_t)label[k];
    }
}

/* ------------------------------------------------------ sliding tile packed */

/* The packed state: tiles[0..size-2] are the non-blank tiles (tile-1) in cell
 * order with the blank removed, tiles[15] is the blank cell.  Packing replaces
 * each tile by the number of smaller tiles to its left; tiles[0] and tiles[1]
 * are dropped (the parity rule restores them).
 * Puzzles/SlidingTilePuzzleOptimized.cpp:79-192. */

static int packed_parity_ok(int width, int size, const int* tiles) {
    int inversions = 0;
    for (int i = 0; i < size - 1; i++)
        for (int j = 0; j < i; j++) inversions += tiles[j] > tiles[i];
    int row = tiles[15] / width;
    if (width % 2 == 0) return (row + inversions) % 2 == 0;
    return inversions % 2 == 0;
}

static void packed_pack_cpu(int size, const int* st, int* out) {
    int t[16] = {0};
    t[15] = st[15];
    for (int i = size - 2; i >= 0; i--) {
        int x = st[i];
        for (int j = i + 1; j < size - 1; j++)
            if (st[i] > st[j]) x--;
        t[i] = x;
    }
    t[0] = 0;
    t[1] = 0;
    memcpy(out, t, sizeof(t));
}

static void packed_unpack_cpu(int width, int size, const int* st, int* out) {
    int t[16] = {0};
    t[15] = st[15];
    for (int i = 0; i < size - 1; i++) {
        int x = st[i];
        for (int j = i + 1; j < size - 1; j++)
            if (x >= st[j]) x++;
        t[i] = x;
    }
    if (!packed_parity_ok(width, size, t)) {
        int a = t[0];
        t[0] = t[1];
        t[1] = a;
    }
    memcpy(out, t, sizeof(t));
}

static void packed_get_index(const int* t, uint32_t* seg, uint32_t* idx) {
    *seg = (uint32_t)(t[14] * 256 + t[13] * 16 + t[12]);
    uint32_t index = 0;
    for (int i = 11; i >= 2; i--) index = index * (uint32_t)(i + 1) + (uint32_t)t[i];
    *idx = index * 16u + (uint32_t)t[15];
}

static void packed_from_index(uint32_t seg, uint32_t index, int* t) {
    memset(t, 0, 16 * sizeof(int));
    t[14] = (int)(seg / 256);
    t[13] = (int)((seg / 16) % 16);
    t[12] = (int)(seg % 16);
    t[15] = (int)(index % 16);
    index /= 16;
    for (int i = 2; i < 12; i++) {
        t[i] = (int)(index % (uint32_t)(i + 1));
        index /= (uint32_t)(i + 1);
    }
}

int64_t or_packed_from_standard(int width, int height, const int* cells) {
    const int size = width * height;
    int st[16] = {0};
    int blank = -1, pos = 0;
    for (int i = 0; i < size; i++) {
        if (cells[i] == 0) {
            blank = i;
            continue;
        }
        st[pos++] = cells[i] - 1;
    }
    if (blank < 0) return -1;
    st[15] = blank;
    if (!packed_parity_ok(width, size, st)) return -1;
    int t[16];
    packed_pack_cpu(size, st, t);
    uint32_t seg, idx;
    packed_get_index(t, &seg, &idx);
    return (int64_t)(((uint64_t)seg << 32) | idx);
}

void or_packed_to_standard(int width, int height, uint64_t index, int* cells) {
    const int size = width * height;
    int t[16], st[16];
    packed_from_index((uint32_t)(index >> 32), (uint32_t)index, t);
    packed_unpack_cpu(width, size, t, st);
    int blank = st[15], pos = 0;
    for (int i = 0; i < size; i++) cells[i] = (i == blank) ? 0 : st[pos++] + 1;
}

/* Kernel-side helpers of kernel_sliding_tile_optimized
 * (Puzzles/SlidingPuzzleGpu.cu:108-207), restated separately from the CPU forms
 * above so the tests can check one against the other. */
static void gk_unpack(int width, int size, int* arr) {
    int inv_even = 1;
    for (int i = 1; i < size - 1; i++)
        for (int j = 0; j < i; j++) {
            int inv = arr[j] >= arr[i];
            arr[j] += inv;
            inv_even ^= inv;
        }
    int row_even = ((arr[15] / width) & 1) == 0;
    int restore = (width % 2 == 0) ? ((!inv_even) == row_even) : !inv_even;
    if (restore) {
        int t = arr[0];
        arr[0] = arr[1];
        arr[1] = t;
    }
}

static void gk_pack(int size, int* arr) {
    arr[0] = 0;
    arr[1] = 0;
    for (int i = size - 2; i > 2; i--)
        for (int j = 2; j < i; j++) arr[j] -= arr[j] >= arr[i];
}

static uint64_t gk_index(const int* arr) {
    uint64_t seg = (uint64_t)(arr[12] | (arr[13] << 4) | (arr[14] << 8));
    uint32_t index = 0;
    for (int i = 11; i >= 2; i--) index = index * (uint32_t)(i + 1) + (uint32_t)arr[i];
    return (seg << 32) | (uint64_t)(index * 16u + (uint32_t)arr[15]);
}

static void gk_rotate_up(int width, int* arr) {
    int z = arr[15];
    int cur = arr[z - width];
    for (int i = z - width; i < z - 1; i++) arr[i] = arr[i + 1];
    arr[z - 1] = cur;
    arr[15] -= width;
}

static void gk_rotate_dn(int width, int* arr) {
    int z = arr[15];
    int cur = arr[z + width - 1];
    for (int i = z + width - 1; i > z; i--) arr[i] = arr[i - 1];
    arr[z] = cur;
    arr[15] += width;
}

static void packed_slots(int width, int height, uint64_t packed, uint64_t* out) {
    const int size = width * height;
    const int op_bits = (int)(packed & 15);
    const uint64_t index = packed >> 4;
    int arr[16];
    packed_from_index((uint32_t)(index >> 32), (uint32_t)index, arr);
    gk_unpack(width, size, arr);
    const int blank = arr[15];
    out[0] = out[1] = out[2] = out[3] = OR_INVALID_INDEX;
    if (blank >= width && !(op_bits & (1 << OP_UP))) {
        int a2[16];
        memcpy(a2, arr, sizeof(a2));
        gk_rotate_up(width, a2);
        gk_pack(size, a2);
        out[0] = (gk_index(a2) << 4) | OP_DOWN;
    }
    if (blank < size - width && !(op_bits & (1 << OP_DOWN))) {
        int a2[16];
        memcpy(a2, arr, sizeof(a2));
        gk_rotate_dn(width, a2);
        gk_pack(size, a2);
        out[1] = (gk_index(a2) << 4) | OP_UP;
    }
    if (blank % width > 0 && !(op_bits & (1 << OP_LEFT))) out[2] = ((index - 1) << 4) | OP_RIGHT;
    if (blank % width < width - 1 && !(op_bits & (1 << OP_RIGHT))) out[3] = ((index + 1) << 4) | OP_LEFT;
}

/* ------------------------------------------------------------ 4-peg Hanoi */

/* Linked-list state of Puzzles/FourPegHanoiSimple.cpp:12-79 and
 * Puzzles/HanoiTowersGPU.cu:31-131: next[] links disks top to bottom. */
typedef struct {
    int next[32];
    int top[4];
    int bottom[4];
} hstate;

static void hs_add(hstate* s, int peg, int disk) {
    if (s->top[peg] == -1) s->top[peg] = disk;
    else s->next[s->bottom[peg]] = disk;
    s->bottom[peg] = disk;
}

static void hs_from_index(int size, uint64_t index, hstate* s) {
    for (int i = 0; i < 32; i++) s->next[i] = -1;
    for (int i = 0; i < 4; i++) s->top[i] = s->bottom[i] = -1;
    for (int d = 0; d < size; d++) hs_add(s, (int)((index >> (2 * d)) & 3), d);
}

static uint64_t hs_to_index(const hstate* s) {
    uint64_t index = 0;
    for (int peg = 0; peg < 4; peg++)
        for (int d = s->top[peg]; d >= 0; d = s->next[d]) index |= (uint64_t)peg << (2 * d);
    return index;
}

static int hs_can_move(const hstate* s, int src, int dst) {
    if (src == dst) return 0;
    int sd = s->top[src];
    if (sd == -1) return 0;
    int dd = s->top[dst];
    if (dd != -1 && dd < sd) return 0;
    return 1;
}

static int hs_move(hstate* s, int src, int dst) {
    if (!hs_can_move(s, src, dst)) return 0;
    int sd = s->top[src], dd = s->top[dst];
    s->top[src] = s->next[sd];
    if (s->top[src] == -1) s->bottom[src] = -1;
    if (dd == -1) {
        s->top[dst] = s->bottom[dst] = sd;
        s->next[sd] = -1;
    } else {
        s->top[dst] = sd;
        s->next[sd] = dd;
    }
    return 1;
}

/* Sort pegs 1..3 by bottom disk, largest first (three compare-swaps), and
 * return the new position of dst. */
static int hs_restore_symmetry(hstate* s, int dst) {
    int pegs[4] = {0, 1, 2, 3};
    const int order[3][2] = {{2, 3}, {1, 2}, {2, 3}};
    for (int k = 0; k < 3; k++) {
        int i = order[k][0], j = order[k][1];
        if (s->bottom[i] < s->bottom[j]) {
            int t;
            t = s->top[i]; s->top[i] = s->top[j]; s->top[j] = t;
            t = s->bottom[i]; s->bottom[i] = s->bottom[j]; s->bottom[j] = t;
            t = pegs[i]; pegs[i] = pegs[j]; pegs[j] = t;
        }
    }
    int inv[4];
    for (int k = 0; k < 4; k++) inv[pegs[k]] = k;
    return inv[dst];
}

/* kernel_hanoi_towers_expand (Puzzles/HanoiTowersGPU.cu:114-144): six slots for
 * peg pairs (0,1),(0,2),(0,3),(1,2),(1,3),(2,3); the smaller top disk moves;
 * the label is the destination peg (after symmetry remapping). */
static void hanoi_slots_linked(int size, int sym, uint64_t packed, uint64_t* out) {
    const int op_bits = (int)(packed & 15);
    const uint64_t index = packed >> 4;
    hstate st;
    hs_from_index(size, index, &st);
    int pos = 0;
    for (int a = 0; a < 4; a++)
        for (int b = a + 1; b < 4; b++) {
            uint64_t r = OR_INVALID_INDEX;
            hstate s2 = st;
            int p1 = a, p2 = b;
            if (!hs_can_move(&s2, p1, p2)) { p1 = b; p2 = a; }
            if (!(op_bits & (1 << p1)) && hs_move(&s2, p1, p2)) {
                if (sym) p2 = hs_restore_symmetry(&s2, p2);
                r = (hs_to_index(&s2) << 4) | (uint64_t)p2;
            }
            out[pos++] = r;
        }
}

/* Bit-plane state of Puzzles/HanoiTowersGPU.cu:167-245 (FPStateOpt): pegs[p]
 * has bit 2d set when disk d is on peg p; top = lowest set bit, 255 if empty. */
typedef struct {
    uint64_t pegs[4];
    int top[4];
} hopt;

static void ho_finish(hopt* s) {
    for (int p = 0; p < 4; p++) s->top[p] = s->pegs[p] ? __builtin_ctzll(s->pegs[p]) : 255;
}

static void ho_from_index(int size, uint64_t index, hopt* s) {
    const uint64_t size_mask = (size >= 32) ? ~0ull : ((1ull << (2 * size)) - 1ull);
    const uint64_t bm = 0x5555555555555555ull & size_mask;
    uint64_t p0 = index >> 1, p1 = index;
    s->pegs[0] = (~p0 & ~p1) & bm;
    s->pegs[1] = (~p0 & p1) & bm;
    s->pegs[2] = (p0 & ~p1) & bm;
    s->pegs[3] = (p0 & p1) & bm;
    ho_finish(s);
}

static uint64_t ho_to_index(const hopt* s) {
    return s->pegs[1] | (s->pegs[2] << 1) | (s->pegs[3] | (s->pegs[3] << 1));
}

static int ho_can_move(const hopt* s, int src, int dst) { return s->top[src] < s->top[dst]; }

static int ho_move(hopt* s, int src, int dst) {
    if (!ho_can_move(s, src, dst)) return 0;
    s->pegs[src] &= ~(1ull << s->top[src]);
    s->pegs[dst] |= 1ull << s->top[src];
    return 1;
}

static int ho_restore_symmetry(hopt* s, int dst) {
    int bottom[4];
    for (int p = 0; p < 4; p++) bottom[p] = s->pegs[p] ? 63 - __builtin_clzll(s->pegs[p]) : -1;
    int pegs[4] = {0, 1, 2, 3};
    const int order[3][2] = {{2, 3}, {1, 2}, {2, 3}};
    for (int k = 0; k < 3; k++) {
        int i = order[k][0], j = order[k][1];
        if (bottom[i] < bottom[j]) {
            int t;
            uint64_t u;
            t = s->top[i]; s->top[i] = s->top[j]; s->top[j] = t;
            t = bottom[i]; bottom[i] = bottom[j]; bottom[j] = t;
            u = s->pegs[i]; s->pegs[i] = s->pegs[j]; s->pegs[j] = u;
            t = pegs[i]; pegs[i] = pegs[j]; pegs[j] = t;
        }
    }
    int inv[4];
    for (int k = 0; k < 4; k++) inv[pegs[k]] = k;
    return inv[dst];
}

/* kernel_hanoi_towers_expand_optimized (Puzzles/HanoiTowersGPU.cu:250-281). */
static void hanoi_slots_opt(int size, int sym, uint64_t packed, uint64_t* out) {
    const int op_bits = (int)(packed & 15);
    const uint64_t index = packed >> 4;
    hopt st;
    ho_from_index(size, index, &st);
    int pos = 0;
    for (int a = 0; a < 4; a++)
        for (int b = a + 1; b < 4; b++) {
            uint64_t r = OR_INVALID_INDEX;
            hopt s2 = st;
            int p1 = a, p2 = b;
            if (!ho_can_move(&s2, p1, p2)) { p1 = b; p2 = a; }
            if (!(op_bits & (1 << p1)) && ho_move(&s2, p1, p2)) {
                if (sym) p2 = ho_restore_symmetry(&s2, p2);
                r = (ho_to_index(&s2) << 4) | (uint64_t)p2;
            }
            out[pos++] = r;
        }
}

int or_hanoi_expand_cpu_opt(int disks, int symmetry, uint64_t index, int op_bits,
                            uint64_t* children, int32_t* ops) {
    hopt st;
    ho_from_index(disks, index, &st);
    int n = 0;
    for (int a = 0; a < 4; a++)
        for (int b = a + 1; b < 4; b++) {
            hopt s2 = st;
            int p1 = a, p2 = b;
            if (!ho_can_move(&s2, p1, p2)) { p1 = b; p2 = a; }
            int dst_was_empty = s2.pegs[p2] == 0;
            if (ho_move(&s2, p1, p2)) {
                int src_is_empty = s2.pegs[p1] == 0;
                if (symmetry && (dst_was_empty || src_is_empty)) p2 = ho_restore_symmetry(&s2, p2);
                if (!(op_bits & (1 << p1))) {
                    children[n] = ho_to_index(&s2);
                    ops[n] = p2;
                    n++;
                }
            }
        }
    return n;
}

/* ------------------------------------------------------ generic interface */

static uint64_t factorial(int n) {
    uint64_t r = 1;
    for (int i = 2; i <= n; i++) r *= (uint64_t)i;
    return r;
}

int or_operators_count(const or_puzzle* p) {
    switch (p->kind) {
    case OR_PANCAKE: return p->size - 1;
    case OR_SLIDING:
    case OR_SLIDING_PACKED:
    case OR_HANOI4:
    case OR_HANOI4_OPT: return 4;
    }
    return -1;
}

int or_branching_factor(const or_puzzle* p) {
    switch (p->kind) {
    case OR_PANCAKE: return p->size - 1;
    case OR_SLIDING:
    case OR_SLIDING_PACKED: return 4;
    case OR_HANOI4:
    case OR_HANOI4_OPT: return 6;
    }
    return -1;
}

uint64_t or_indexes_count(const or_puzzle* p) {
    switch (p->kind) {
    case OR_PANCAKE: {
        /* Puzzles/PancakeOptimized.cpp:94-105 (n = 12 gives 2^29, not 12!) */
        if (p->size < 12) return factorial(p->size);
        uint64_t r = 1;
        for (int i = 13; i <= p->size; i++) r *= (uint64_t)i;
        return r << 29;
    }
    case OR_SLIDING: return factorial(p->size * p->height);
    case OR_SLIDING_PACKED: {
        /* Puzzles/SlidingTilePuzzleOptimized.cpp:12-26, 271-277 */
        int size = p->size * p->height;
        uint64_t segs = size == 16 ? 0xEDC + 1 : size == 15 ? 0xDC + 1 : size == 14 ? 0xC + 1 : 1;
        if (segs > 1) return segs << 32;
        uint64_t m = 16;
        for (int i = 3; i < size; i++) m *= (uint64_t)i;
        return m;
    }
    case OR_HANOI4:
    case OR_HANOI4_OPT: return 1ull << (2 * p->size);
    }
    return 0;
}

int or_has_odd_cycles(const or_puzzle* p) {
    return p->kind == OR_PANCAKE || p->kind == OR_HANOI4 || p->kind == OR_HANOI4_OPT;
}

int or_slots_per_parent(const or_puzzle* p, int mode) {
    if (p->kind == OR_PANCAKE) {
        if (mode == OR_EXPAND_FULL) return p->size - 1;
        if (p->size <= 12) return -1;
        return mode == OR_EXPAND_IN_SEGMENT ? 11 : p->size - 12;
    }
    if (mode != OR_EXPAND_FULL) return -1;
    return or_branching_factor(p);
}

int or_expand_slots(const or_puzzle* p, int mode, const uint64_t* in, uint64_t count, uint64_t* out) {
    const int s = or_slots_per_parent(p, mode);
    if (s < 0) return -1;
    for (uint64_t i = 0; i < count; i++) {
        uint64_t* o = out + i * (uint64_t)s;
        switch (p->kind) {
        case OR_PANCAKE: pancake_slots(p->size, (p->flags & OR_FLAG_INV_INDEX) != 0, mode, in[i], o); break;
        case OR_SLIDING: sliding_slots(p->size, p->size * p->height, in[i], o); break;
        case OR_SLIDING_PACKED: packed_slots(p->size, p->height, in[i], o); break;
        case OR_HANOI4: hanoi_slots_linked(p->size, p->flags & OR_FLAG_SYMMETRY, in[i], o); break;
        case OR_HANOI4_OPT: hanoi_slots_opt(p->size, p->flags & OR_FLAG_SYMMETRY, in[i], o); break;
        default: return -1;
        }
    }
    return s;
}

int or_expand_one(const or_puzzle* p, int mode, uint64_t index, int op_bits,
                  uint64_t* children, int32_t* ops) {
    uint64_t slots[16];
    const int ops_count = or_operators_count(p);
    uint64_t packed = (index << ops_count) | (uint64_t)op_bits;
    int s = or_expand_slots(p, mode, &packed, 1, slots);
    if (s < 0) return -1;
    int n = 0;
    const uint64_t mask = (1ull << ops_count) - 1ull;
    for (int k = 0; k < s; k++) {
        if (slots[k] == OR_INVALID_INDEX) continue;
        children[n] = slots[k] >> ops_count;
        ops[n] = (int32_t)(slots[k] & mask);
        n++;
    }
    return n;
}

void or_combine_index_opbits(int ops, const uint64_t* idx, const int32_t* op_bits, uint64_t count, uint64_t* out) {
    for (uint64_t i = 0; i < count; i++) out[i] = (idx[i] << ops) | (uint64_t)(uint32_t)op_bits[i];
}

void or_split_index_ops(int ops, const uint64_t* packed, uint64_t count, uint64_t* idx, int32_t* op) {
    const uint64_t mask = (1ull << ops) - 1ull;
    for (uint64_t i = 0; i < count; i++) {
        op[i] = (int32_t)(packed[i] & mask);
        idx[i] = (uint64_t)((int64_t)packed[i] >> ops); /* arithmetic shift keeps ~0 */
    }
}

/* ------------------------------------------------------ text formats */

static int read_ints(const char* text, int* vals, int cap) {
    int n = 0;
    const char* s = text;
    while (*s && n < cap) {
        char* end;
        long v = strtol(s, &end, 10);
        if (end == s) break;
        vals[n++] = (int)v;
        s = end;
    }
    return n;
}

uint64_t or_parse(const or_puzzle* p, const char* text) {
    int vals[64];
    int n = read_ints(text, vals, 64);
    switch (p->kind) {
    case OR_PANCAKE:
        if (n < p->size) return OR_INVALID_INDEX;
        return or_pancake_rank(vals, p->size);
    case OR_SLIDING:
        if (n < p->size * p->height) return OR_INVALID_INDEX;
        return or_perm_rank(vals, p->size * p->height);
    case OR_SLIDING_PACKED: {
        if (n < p->size * p->height) return OR_INVALID_INDEX;
        int64_t r = or_packed_from_standard(p->size, p->height, vals);
        return r < 0 ? OR_INVALID_INDEX : (uint64_t)r;
    }
    case OR_HANOI4:
    case OR_HANOI4_OPT: {
        /* Puzzles/FourPegHanoiSimple.cpp:126-155: disks 1..n per peg top to
         * bottom, 0 separates pegs. */
        uint64_t index = 0;
        int peg = 0;
        for (int i = 0; i < n; i++) {
            if (vals[i] == 0) peg++;
            else index |= (uint64_t)peg << (2 * (vals[i] - 1));
        }
        return index;
    }
    }
    return OR_INVALID_INDEX;
}

int or_to_string(const or_puzzle* p, uint64_t index, char* buf, int cap) {
    char tmp[512];
    int len = 0;
    tmp[0] = 0;
    switch (p->kind) {
    case OR_PANCAKE:
    case OR_SLIDING:
    case OR_SLIDING_PACKED: {
        int arr[16];
        int size = p->kind == OR_PANCAKE ? p->size : p->size * p->height;
        if (p->kind == OR_PANCAKE) or_pancake_unrank(index, arr, size);
        else if (p->kind == OR_SLIDING) or_perm_unrank(index, arr, size);
        else or_packed_to_standard(p->size, p->height, index, arr);
        for (int i = 0; i < size; i++) len += snprintf(tmp + len, sizeof(tmp) - len, i ? " %d" : "%d", arr[i]);
        break;
    }
    case OR_HANOI4:
    case OR_HANOI4_OPT: {
        /* Puzzles/FourPegHanoiSimple.cpp:114-124 */
        for (int peg = 0; peg < 4; peg++) {
            for (int d = 0; d < p->size; d++)
                if ((int)((index >> (2 * d)) & 3) == peg) len += snprintf(tmp + len, sizeof(tmp) - len, "%d ", d + 1);
            len += snprintf(tmp + len, sizeof(tmp) - len, "0 ");
        }
        len -= 3;
        tmp[len] = 0;
        break;
    }
    default: return -1;
    }
    if (len + 1 > cap) return -1;
    memcpy(buf, tmp, (size_t)len + 1);
    return len;
}
