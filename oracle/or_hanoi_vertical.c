/*
 * or_hanoi_vertical.c -- the HanoiTowers "vertical" solver restated: a
 * symmetry-reduced three-bit BFS of the 4-peg Towers of Hanoi over segments of
 * 2^seg_bits states (2^32 in the reference), with the layer counts restored to
 * the unreduced puzzle by PegsMult.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Reference:
 *   HanoiTowers/HanoiTowers.h:37-55        InSegmentOnly, PegsMult
 *   HanoiTowers/HanoiTowers.cpp:10-133     FPState (pegs as bit masks, move, restore_symmetry)
 *   HanoiTowers/HanoiTowers.cpp:152-172    Expand(index)
 *   HanoiTowers/HanoiTowers.cpp:174-198    ExpandCrossSegment
 *   HanoiTowers/HanoiTowers.cpp:200-300    ExpandInSegment, ExpandInSegmentWithoutSmallest
 *   HanoiTowers/HanoiExpandGPU.cu:98-229   the GPU kernels' slot form (6 or 3 slots per parent,
 *                                          the parent's own id where a child leaves the segment)
 *   HanoiTowers/HanoiSearch.cpp:46-175     HanoiSolver::SetInitialNode, Expand(segment)
 *   HanoiTowers/HanoiSearch.cpp:215-290    HanoiSearch: the layer loop, result[0] = 1
 *
 * A state index holds disk d's peg in bits 2d..2d+1.  The reduction keeps only
 * states whose pegs 1..3 are ordered by their bottom (largest) disk, largest
 * first, empty pegs last.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

#define HV_MAX_DISKS 24
#define HV_BIT_MASK 0x5555555555555555ull

typedef struct {
    uint64_t pegs[4]; /* bit 2d set: disk d is on the peg */
    int top[4];       /* bit position of the top disk, 255 when empty */
} hv_state;

static int hv_lowbit(uint64_t x) { return x ? __builtin_ctzll(x) : 255; }
static int hv_highbit(uint64_t x) { return x ? 63 - __builtin_clzll(x) : -1; }

/* FPState::from_index (HanoiTowers.cpp:15-49) */
static void hv_from_index(int disks, uint64_t index, hv_state* s) {
    const uint64_t size_mask = disks >= 32 ? ~0ull : (1ull << (2 * disks)) - 1;
    const uint64_t m = HV_BIT_MASK & size_mask;
    const uint64_t p0 = index >> 1, p1 = index;
    s->pegs[0] = (~p0 & ~p1) & m;
    s->pegs[1] = (~p0 & p1) & m;
    s->pegs[2] = (p0 & ~p1) & m;
    s->pegs[3] = (p0 & p1) & m;
    for (int p = 0; p < 4; ++p) s->top[p] = hv_lowbit(s->pegs[p]);
}

/* FPState::to_index (:65-67) */
static uint64_t hv_to_index(const hv_state* s) {
    return s->pegs[1] | (s->pegs[2] << 1) | ((s->pegs[3] << 1) | s->pegs[3]);
}

/* FPState::move (:79-86): the smaller of the two top disks changes peg */
static void hv_move(hv_state* s, int a, int b) {
    const int disk = s->top[a] < s->top[b] ? s->top[a] : s->top[b];
    if (disk == 255) return;
    const uint64_t bit = 1ull << disk;
    s->pegs[a] ^= bit;
    s->pegs[b] ^= bit;
    s->top[a] = hv_lowbit(s->pegs[a]);
    s->top[b] = hv_lowbit(s->pegs[b]);
}

/* FPState::restore_symmetry (:88-106): order pegs 1..3 by bottom disk */
static void hv_restore(hv_state* s) {
    int bottom[4];
    for (int p = 0; p < 4; ++p) bottom[p] = hv_highbit(s->pegs[p]);
    static const int pairs[3][2] = {{2, 3}, {1, 2}, {2, 3}};
    for (int k = 0; k < 3; ++k) {
        const int i = pairs[k][0], j = pairs[k][1];
        if (bottom[i] < bottom[j]) {
            int t = bottom[i]; bottom[i] = bottom[j]; bottom[j] = t;
            t = s->top[i]; s->top[i] = s->top[j]; s->top[j] = t;
            uint64_t u = s->pegs[i]; s->pegs[i] = s->pegs[j]; s->pegs[j] = u;
        }
    }
}

/* one move of Expand / ExpandInSegment's symmetric branch: restore only when
 * one of the two pegs is or becomes empty (HanoiTowers.cpp:156-165) */
static uint64_t hv_child(const hv_state* s, int a, int b) {
    hv_state t = *s;
    const int was_empty = t.pegs[a] == 0 || t.pegs[b] == 0;
    hv_move(&t, a, b);
    const int is_empty = t.pegs[a] == 0 || t.pegs[b] == 0;
    if (was_empty || is_empty) hv_restore(&t);
    return hv_to_index(&t);
}

static const int HV_PAIRS[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
/* disksWithoutSmallest (:255-260): the pegs other than the smallest disk's */
static const int HV_OTHER[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};

int or_hv_pegs_mult(uint64_t index) {
    /* HanoiTowers.h:48-55 */
    const uint64_t p0 = index >> 1, p1 = index;
    const uint64_t pegs1 = (~p0 & p1) & HV_BIT_MASK;
    const uint64_t pegs2 = (p0 & ~p1) & HV_BIT_MASK;
    return 6 - (pegs2 == 0) * 3 - (pegs1 == 0);
}

int or_hv_in_segment_only(uint64_t index, int seg_bits) {
    /* HanoiTowers.h:37-46: at least three pegs hold one of the in-segment
     * disks, so no disk above them can move */
    const uint64_t low = seg_bits >= 64 ? ~0ull : (1ull << seg_bits) - 1;
    const uint64_t m = HV_BIT_MASK & low;
    const uint64_t p0 = index >> 1, p1 = index;
    const int z0 = ((~p0 & ~p1) & m) != 0, z1 = ((~p0 & p1) & m) != 0;
    const int z2 = ((p0 & ~p1) & m) != 0, z3 = ((p0 & p1) & m) != 0;
    return z0 + z1 + z2 + z3 >= 3;
}

void or_hv_expand(int disks, uint64_t index, uint64_t* out6) {
    /* HanoiTowers.cpp:152-172 */
    hv_state s;
    hv_from_index(disks, index, &s);
    for (int k = 0; k < 6; ++k) out6[k] = hv_child(&s, HV_PAIRS[k][0], HV_PAIRS[k][1]);
}

int or_hv_expand_cross_segment(int disks, int seg_bits, uint64_t segment, uint32_t index, uint64_t* out) {
    /* HanoiTowers.cpp:174-198: the children outside the segment, in move order */
    uint64_t c[6];
    or_hv_expand(disks, (segment << seg_bits) | index, c);
    int n = 0;
    for (int k = 0; k < 6; ++k)
        if ((c[k] >> seg_bits) != segment) out[n++] = c[k];
    return n;
}

void or_hv_expand_in_segment_slots(int disks, uint32_t segment, const uint32_t* idx, uint64_t count,
                                   int without_smallest, uint32_t* out) {
    /* HanoiExpandGPU.cu:98-229 with 2^32-state segments: slot k of parent i at
     * out[i*S + k] (S = 6, or 3 without the smallest disk's moves); a child in
     * another segment is written as the parent's own id */
    const uint64_t base = (uint64_t)segment << 32;
    for (uint64_t i = 0; i < count; ++i) {
        const uint64_t index = base | idx[i];
        hv_state s;
        hv_from_index(disks, index, &s);
        if (!without_smallest) {
            for (int k = 0; k < 6; ++k) {
                const uint64_t c = hv_child(&s, HV_PAIRS[k][0], HV_PAIRS[k][1]);
                out[i * 6 + k] = (c >> 32) == segment ? (uint32_t)c : idx[i];
            }
        } else {
            const int* o = HV_OTHER[index & 3];
            const int pr[3][2] = {{o[0], o[1]}, {o[0], o[2]}, {o[1], o[2]}};
            for (int k = 0; k < 3; ++k) {
                const uint64_t c = hv_child(&s, pr[k][0], pr[k][1]);
                out[i * 3 + k] = (c >> 32) == segment ? (uint32_t)c : idx[i];
            }
        }
    }
}

/* ---- the search (HanoiSearch.cpp) with in-memory stores ---- */

typedef struct {
    uint32_t* v;
    size_t n, cap;
} hv_list;

static int hv_push(hv_list* l, uint32_t x) {
    if (l->n == l->cap) {
        size_t c = l->cap ? l->cap * 2 : 1024;
        uint32_t* p = (uint32_t*)realloc(l->v, c * sizeof(uint32_t));
        if (!p) return -1;
        l->v = p;
        l->cap = c;
    }
    l->v[l->n++] = x;
    return 0;
}

/* IndexedBitArray (Common/BitArray.h:201-281): a bit array with an occupancy
 * flag per 1024 words, so scans and clears visit only touched blocks */
typedef struct {
    uint64_t* w;
    uint8_t* used;
    uint64_t nwords, nblocks;
} hv_bits;

static int hv_bits_init(hv_bits* b, uint64_t nbits) {
    b->nwords = (nbits + 63) / 64;
    b->nblocks = (b->nwords + 1023) / 1024;
    b->w = (uint64_t*)calloc(b->nwords, 8);
    b->used = (uint8_t*)calloc(b->nblocks, 1);
    return b->w && b->used ? 0 : -1;
}
static void hv_bits_free(hv_bits* b) {
    free(b->w);
    free(b->used);
}
static void hv_set(hv_bits* b, uint64_t i) {
    b->w[i >> 6] |= 1ull << (i & 63);
    b->used[(i >> 6) >> 10] = 1;
}
static void hv_clear_bit(hv_bits* b, uint64_t i) { b->w[i >> 6] &= ~(1ull << (i & 63)); }
static void hv_clear_all(hv_bits* b) {
    for (uint64_t k = 0; k < b->nblocks; ++k) {
        if (!b->used[k]) continue;
        const uint64_t w0 = k * 1024, w1 = w0 + 1024 < b->nwords ? w0 + 1024 : b->nwords;
        memset(b->w + w0, 0, (w1 - w0) * 8);
        b->used[k] = 0;
    }
}

int or_hv_search(int disks, uint64_t start, int seg_bits, int max_layers, uint64_t* restored, uint64_t* reduced,
                 uint64_t* hashes) {
    if (disks < 3 || disks > HV_MAX_DISKS || seg_bits < 6 || seg_bits > 32 || (seg_bits & 1) || max_layers < 1)
        return -1;
    const int total_bits = 2 * disks;
    const int sb = seg_bits < total_bits ? seg_bits : total_bits;
    if (total_bits - sb > 24) return -1;
    const uint64_t nseg = 1ull << (total_bits - sb);
    const uint64_t seg_states = 1ull << sb;
    if (start >> total_bits) return -1;

    /* stores: [0] old, [1] cur, [2] new frontier; [3] current and [4] next cross-segment marks */
    hv_list* st[5];
    for (int k = 0; k < 5; ++k) {
        st[k] = (hv_list*)calloc(nseg, sizeof(hv_list));
        if (!st[k]) return -1;
    }
    hv_bits next, curb;
    if (hv_bits_init(&next, seg_states) || hv_bits_init(&curb, seg_states)) return -1;
    hv_list xs = {0, 0, 0};
    int layers = 0, rc = 0;

    /* SetInitialNode (HanoiSearch.cpp:46-69): the start in its segment's new
     * frontier, its cross-segment children in the next marks */
    {
        const uint64_t s0 = start >> sb;
        const uint32_t i0 = (uint32_t)(start & (seg_states - 1));
        if (hv_push(&st[2][s0], i0)) rc = -1;
        uint64_t c[6];
        const int n = or_hv_expand_cross_segment(disks, sb, s0, i0, c);
        for (int k = 0; k < n && !rc; ++k)
            if (hv_push(&st[4][c[k] >> sb], (uint32_t)(c[k] & (seg_states - 1)))) rc = -1;
    }
    restored[0] = 1;
    reduced[0] = 1;
    if (hashes) hashes[0] = or_splitmix64(or_splitmix64(start >> 6) ^ (1ull << (start & 63)));
    layers = 1;

    for (;;) {
        /* fnSwapStores (:232-238) */
        for (uint64_t s = 0; s < nseg; ++s) {
            st[0][s].n = 0;
            hv_list t = st[0][s]; st[0][s] = st[1][s]; st[1][s] = st[2][s]; st[2][s] = t;
            st[3][s].n = 0;
            t = st[3][s]; st[3][s] = st[4][s]; st[4][s] = t;
        }
        if (layers >= max_layers || rc) break;
        uint64_t total = 0, total_restored = 0, hash = 0;
        for (uint64_t seg = 0; seg < nseg && !rc; ++seg) {
            /* HanoiSolver::Expand (HanoiSearch.cpp:71-175) */
            const int full_sym = or_hv_pegs_mult(seg) == 6;
            int has_data = 0;
            for (size_t i = 0; i < st[3][seg].n; ++i) {
                hv_set(&next, st[3][seg].v[i]);
                has_data = 1;
            }
            const hv_list* cur = &st[1][seg];
            if (cur->n) has_data = 1;
            const uint64_t base = seg << sb;
            for (size_t i = 0; i < cur->n; ++i) {
                const uint32_t id = cur->v[i];
                hv_set(&curb, id);
                hv_state s;
                hv_from_index(disks, base | id, &s);
                if (full_sym) {
                    /* SetNextFourBits of the id with the smallest disk's peg
                     * cleared, then the moves among the other three pegs */
                    const uint32_t b4 = id & ~3u;
                    for (uint32_t q = 0; q < 4; ++q) hv_set(&next, b4 | q);
                    const int* o = HV_OTHER[id & 3];
                    const int pr[3][2] = {{o[0], o[1]}, {o[0], o[2]}, {o[1], o[2]}};
                    for (int k = 0; k < 3; ++k) {
                        const uint64_t c = hv_child(&s, pr[k][0], pr[k][1]);
                        if ((c >> sb) == seg) hv_set(&next, c & (seg_states - 1));
                    }
                } else {
                    for (int k = 0; k < 6; ++k) {
                        const uint64_t c = hv_child(&s, HV_PAIRS[k][0], HV_PAIRS[k][1]);
                        if ((c >> sb) == seg) hv_set(&next, c & (seg_states - 1));
                    }
                }
            }
            if (!has_data) continue;
            for (size_t i = 0; i < st[0][seg].n; ++i) hv_clear_bit(&next, st[0][seg].v[i]);
            /* ScanBitsAndClearWithExcl(CurArray) in index order */
            xs.n = 0;
            for (uint64_t k = 0; k < next.nblocks && !rc; ++k) {
                if (!next.used[k]) continue;
                const uint64_t w0 = k * 1024, w1 = w0 + 1024 < next.nwords ? w0 + 1024 : next.nwords;
                for (uint64_t j = w0; j < w1; ++j) {
                    uint64_t w = next.w[j] & ~curb.w[j];
                    next.w[j] = 0;
                    if (!w) continue;
                    const uint64_t gw = (base >> 6) + j;
                    hash += or_splitmix64(or_splitmix64(gw) ^ w);
                    while (w) {
                        const int b = __builtin_ctzll(w);
                        w &= w - 1;
                        const uint32_t id = (uint32_t)(j * 64 + b);
                        ++total;
                        total_restored += full_sym ? 6 : (uint64_t)or_hv_pegs_mult(base | id);
                        if (!or_hv_in_segment_only(id, sb) && hv_push(&xs, id)) rc = -1;
                        if (hv_push(&st[2][seg], id)) rc = -1;
                    }
                }
                next.used[k] = 0;
            }
            hv_clear_all(&curb);
            /* Expander::ExpandCrossSegment, Mult.Add */
            for (size_t i = 0; i < xs.n && !rc; ++i) {
                uint64_t c[6];
                const int n = or_hv_expand_cross_segment(disks, sb, seg, xs.v[i], c);
                for (int k = 0; k < n && !rc; ++k)
                    if (hv_push(&st[4][c[k] >> sb], (uint32_t)(c[k] & (seg_states - 1)))) rc = -1;
            }
        }
        if (total == 0 || rc) break;
        restored[layers] = total_restored;
        reduced[layers] = total;
        if (hashes) hashes[layers] = hash;
        ++layers;
    }
    for (int k = 0; k < 5; ++k) {
        for (uint64_t s = 0; s < nseg; ++s) free(st[k][s].v);
        free(st[k]);
    }
    free(xs.v);
    hv_bits_free(&next);
    hv_bits_free(&curb);
    return rc ? -1 : layers;
}
