/*
 * or_bfs.c -- CPU restatement of the reference bit arrays and BFS drivers
 * (the oracle; test infrastructure only, see oracle.h).
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

/* Common/BitArray.h:97-125 */
void or_bits_andnot(uint64_t* a, const uint64_t* excl, uint64_t nwords) {
    for (uint64_t i = 0; i < nwords; i++) a[i] &= ~excl[i];
}

uint64_t or_bits_andnot_count(uint64_t* a, const uint64_t* excl, uint64_t nwords) {
    uint64_t r = 0;
    for (uint64_t i = 0; i < nwords; i++) {
        uint64_t v = a[i] & ~excl[i];
        a[i] = v;
        r += (uint64_t)__builtin_popcountll(v);
    }
    return r;
}

uint64_t or_bits_count(const uint64_t* a, uint64_t nwords) {
    uint64_t r = 0;
    for (uint64_t i = 0; i < nwords; i++) r += (uint64_t)__builtin_popcountll(a[i]);
    return r;
}

static inline void bit_set(uint64_t* a, uint64_t i) { a[i >> 6] |= 1ull << (i & 63); }
static inline void bit_clear(uint64_t* a, uint64_t i) { a[i >> 6] &= ~(1ull << (i & 63)); }

/* Expands every set bit of cur in [first, first+span) (ascending, as
 * BitArray::ScanBits at Common/BitArray.h:137-144) and sets each child in next,
 * as InMemoryThreeBitBFS's ExpandBuffer callback does
 * (Common/InMemoryThreeBitBFS.cpp:17-30, Common/Puzzle.h:85-106). */
static int64_t expand_mark(const or_puzzle* p, const uint64_t* cur, uint64_t* next, uint64_t n_index,
                           uint64_t first, uint64_t span, uint64_t* children_out) {
    uint64_t children[16];
    int32_t ops[16];
    uint64_t parents = 0, marks = 0;
    uint64_t end = first + span;
    if (end > n_index) end = n_index;
    for (uint64_t w = first >> 6; w < (end + 63) >> 6; w++) {
        uint64_t val = cur[w];
        while (val) {
            uint64_t i = (w << 6) | (uint64_t)__builtin_ctzll(val);
            val &= val - 1;
            if (i < first || i >= end) continue;
            int n = or_expand_one(p, OR_EXPAND_FULL, i, 0, children, ops);
            if (n < 0) return -1;
            for (int k = 0; k < n; k++) {
                if (children[k] >= n_index) return -1;
                bit_set(next, children[k]);
            }
            parents++;
            marks += (uint64_t)n;
        }
    }
    if (children_out) *children_out = marks;
    return (int64_t)parents;
}

uint64_t or_bfs3_expand_range(const or_puzzle* p, const uint64_t* cur, uint64_t* next,
                              uint64_t first, uint64_t span, uint64_t* children_out) {
    int64_t r = expand_mark(p, cur, next, UINT64_MAX, first, span, children_out);
    return r < 0 ? 0 : (uint64_t)r;
}

/* Common/InMemoryThreeBitBFS.cpp:4-48 */
int or_bfs3(const or_puzzle* p, uint64_t start, uint64_t index_count, int max_layers,
            uint64_t* counts, uint64_t* hashes) {
    const uint64_t n_index = index_count ? index_count : or_indexes_count(p);
    if (start >= n_index || max_layers < 1) return -1;
    const uint64_t nw = (n_index + 63) / 64;
    uint64_t* old = calloc(nw, 8);
    uint64_t* cur = calloc(nw, 8);
    uint64_t* nxt = calloc(nw, 8);
    int layers = -1;
    if (!old || !cur || !nxt) goto done;
    bit_set(cur, start);
    counts[0] = 1;
    if (hashes) hashes[0] = or_bits_hash(cur, nw);
    layers = 1;
    while (layers < max_layers) {
        if (expand_mark(p, cur, nxt, n_index, 0, n_index, NULL) < 0) {
            layers = -1;
            break;
        }
        if (or_has_odd_cycles(p)) or_bits_andnot(nxt, cur, nw);
        uint64_t count = or_bits_andnot_count(nxt, old, nw);
        if (count == 0) break;
        counts[layers] = count;
        if (hashes) hashes[layers] = or_bits_hash(nxt, nw);
        layers++;
        uint64_t* t = old;
        old = cur;
        cur = nxt;
        nxt = t;
        memset(nxt, 0, nw * 8);
    }
done:
    free(old);
    free(cur);
    free(nxt);
    return layers;
}

/* Common/InMemoryTwoBitBFS.cpp:24-69: UNVISITED=0, OLD=1, CUR=2, NEXT=3 with CUR
 * and NEXT swapped after every layer. */
static inline int tb_get(const uint64_t* a, uint64_t i) { return (int)((a[i >> 5] >> (2 * (i & 31))) & 3); }
static inline void tb_set(uint64_t* a, uint64_t i, int v) {
    uint64_t sh = 2 * (i & 31);
    a[i >> 5] = (a[i >> 5] & ~(3ull << sh)) | ((uint64_t)v << sh);
}

int or_bfs2(const or_puzzle* p, uint64_t start, uint64_t index_count, int max_layers,
            uint64_t* counts, uint64_t* hashes) {
    const uint64_t n_index = index_count ? index_count : or_indexes_count(p);
    if (start >= n_index || max_layers < 1) return -1;
    const uint64_t nw = (n_index + 31) / 32;
    uint64_t* arr = calloc(nw, 8);
    if (!arr) return -1;
    int UNVISITED = 0, OLD = 1, CUR = 2, NEXT = 3;
    tb_set(arr, start, CUR);
    counts[0] = 1;
    if (hashes) hashes[0] = or_bits_hash(arr, nw);
    int layers = 1;
    uint64_t children[16];
    int32_t ops[16];
    while (layers < max_layers) {
        uint64_t count = 0;
        for (uint64_t i = 0; i < n_index; i++) {
            if (tb_get(arr, i) != CUR) continue;
            tb_set(arr, i, OLD);
            int n = or_expand_one(p, OR_EXPAND_FULL, i, 0, children, ops);
            if (n < 0) {
                free(arr);
                return -1;
            }
            for (int k = 0; k < n; k++) {
                if (children[k] >= n_index) {
                    free(arr);
                    return -1;
                }
                if (tb_get(arr, children[k]) == UNVISITED) {
                    count++;
                    tb_set(arr, children[k], NEXT);
                }
            }
        }
        if (count == 0) break;
        counts[layers] = count;
        if (hashes) hashes[layers] = or_bits_hash(arr, nw);
        layers++;
        int t = CUR;
        CUR = NEXT;
        NEXT = t;
    }
    free(arr);
    return layers;
}

/* --------------------------------------------- segmented three-bit BFS */

typedef struct {
    uint32_t* v;
    uint64_t n, cap;
} u32list;

static void ul_add(u32list* l, uint32_t x) {
    if (l->n == l->cap) {
        l->cap = l->cap ? l->cap * 2 : 64;
        l->v = realloc(l->v, l->cap * 4);
    }
    l->v[l->n++] = x;
}

static void ul_free(u32list* l) {
    free(l->v);
    l->v = NULL;
    l->n = l->cap = 0;
}

/* Expand with the reference ExpandHint {SegmentBits, CrossSegment}: the pancake
 * 29-bit split uses the in/cross-segment kernels, everything else the full
 * expansion (Puzzles/PancakeOptimizedGpu.cpp:36-79). */
static int expand_hint(const or_puzzle* p, int seg_bits, int cross, uint64_t index,
                       uint64_t* children, int32_t* ops) {
    int mode = OR_EXPAND_FULL;
    if (p->kind == OR_PANCAKE && seg_bits == 29 && p->size > 12) mode = cross ? OR_EXPAND_CROSS_SEGMENT : OR_EXPAND_IN_SEGMENT;
    return or_expand_one(p, mode, index, 0, children, ops);
}

/* Common/DiskBasedOptThreeBitBFS.cpp:54-144 and :164-252, with each Store
 * replaced by a per-segment in-memory list (order inside a list does not change
 * the bit arrays it is folded into). */
int or_bfs3_segmented(const or_puzzle* p, uint64_t start, int segment_bits, int max_layers,
                      uint64_t* counts) {
    const uint64_t total = or_indexes_count(p);
    uint64_t seg_size = 1ull << segment_bits;
    const uint64_t seg_mask = seg_size - 1;
    const uint64_t nseg = (total + seg_size - 1) / seg_size;
    if (nseg == 1 && seg_size > total) seg_size = total;
    const int nops = or_operators_count(p);
    const int odd = or_has_odd_cycles(p);
    const uint64_t nw = (seg_size + 63) / 64;
    u32list* fr_old = calloc(nseg, sizeof(u32list));
    u32list* fr_cur = calloc(nseg, sizeof(u32list));
    u32list* fr_new = calloc(nseg, sizeof(u32list));
    u32list* x_cur = calloc(nseg * (uint64_t)nops, sizeof(u32list));
    u32list* x_next = calloc(nseg * (uint64_t)nops, sizeof(u32list));
    uint64_t* next_arr = calloc(nw, 8);
    uint64_t* cur_arr = calloc(nw, 8);
    uint64_t children[16];
    int32_t ops[16];
    int layers = -1;
    if (!fr_old || !fr_cur || !fr_new || !x_cur || !x_next || !next_arr || !cur_arr) goto done;

    {   /* SetInitialNode (:34-52) then fnSwapStores */
        uint64_t seg = start >> segment_bits;
        ul_add(&fr_cur[seg], (uint32_t)(start & seg_mask));
        int n = expand_hint(p, segment_bits, 1, start, children, ops);
        for (int k = 0; k < n; k++) {
            uint64_t s = children[k] >> segment_bits;
            if (s == seg) continue;
            ul_add(&x_cur[s * (uint64_t)nops + (uint64_t)ops[k]], (uint32_t)(children[k] & seg_mask));
        }
    }
    counts[0] = 1;
    layers = 1;
    while (layers < max_layers) {
        uint64_t total_count = 0;
        for (uint64_t segment = 0; segment < nseg; segment++) {
            const uint64_t base = segment << segment_bits;
            int has_data = 0;
            for (int op = 0; op < nops; op++) {
                u32list* l = &x_cur[segment * (uint64_t)nops + (uint64_t)op];
                if (l->n == 0) continue;
                has_data = 1;
                if (p->kind == OR_PANCAKE && (p->flags & OR_FLAG_INV_INDEX) && segment_bits == 29 && p->size > 12)
                    or_pancake_xseg_postprocess(p->size, (int)segment, l->v, l->n);
                for (uint64_t i = 0; i < l->n; i++) bit_set(next_arr, l->v[i]);
                ul_free(l);
            }
            u32list* cl = &fr_cur[segment];
            for (uint64_t i = 0; i < cl->n; i++) {
                has_data = 1;
                int n = expand_hint(p, segment_bits, 0, base | cl->v[i], children, ops);
                for (int k = 0; k < n; k++)
                    if ((children[k] >> segment_bits) == segment) bit_set(next_arr, children[k] & seg_mask);
                if (odd) bit_set(cur_arr, cl->v[i]);
            }
            if (!has_data) continue;
            u32list* ol = &fr_old[segment];
            for (uint64_t i = 0; i < ol->n; i++) bit_clear(next_arr, ol->v[i]);
            ul_free(ol);
            for (uint64_t w = 0; w < nw; w++) {
                uint64_t val = next_arr[w];
                if (!val) continue;
                next_arr[w] = 0;
                if (odd) val &= ~cur_arr[w];
                while (val) {
                    uint64_t idx = (w << 6) | (uint64_t)__builtin_ctzll(val);
                    val &= val - 1;
                    total_count++;
                    ul_add(&fr_new[segment], (uint32_t)idx);
                    int n = expand_hint(p, segment_bits, 1, base | idx, children, ops);
                    for (int k = 0; k < n; k++) {
                        uint64_t s = children[k] >> segment_bits;
                        if (s == segment) continue;
                        ul_add(&x_next[s * (uint64_t)nops + (uint64_t)ops[k]], (uint32_t)(children[k] & seg_mask));
                    }
                }
            }
            if (odd) memset(cur_arr, 0, nw * 8);
        }
        if (total_count == 0) break;
        counts[layers++] = total_count;
        /* fnSwapStores (:182-187) */
        for (uint64_t s = 0; s < nseg; s++) ul_free(&fr_old[s]);
        u32list* t = fr_old;
        fr_old = fr_cur;
        fr_cur = fr_new;
        fr_new = t;
        for (uint64_t s = 0; s < nseg * (uint64_t)nops; s++) ul_free(&x_cur[s]);
        t = x_cur;
        x_cur = x_next;
        x_next = t;
    }
done:
    if (fr_old)
        for (uint64_t s = 0; s < nseg; s++) {
            ul_free(&fr_old[s]);
            ul_free(&fr_cur[s]);
            ul_free(&fr_new[s]);
        }
    if (x_cur)
        for (uint64_t s = 0; s < nseg * (uint64_t)nops; s++) {
            ul_free(&x_cur[s]);
            ul_free(&x_next[s]);
        }
    free(fr_old);
    free(fr_cur);
    free(fr_new);
    free(x_cur);
    free(x_next);
    free(next_arr);
    free(cur_arr);
    return layers;
}
