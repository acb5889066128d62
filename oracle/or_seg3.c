/*
 * or_seg3.c -- one segment step of the disk-based three-bit BFS, CPU
 * restatement.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * DB_Opt3BitBFS_Solver::Expand(segment) (Common/DiskBasedOptThreeBitBFS.cpp:54-144)
 * with the disk stores replaced by in-memory lists and streams:
 *   1  cross-segment marks of the previous step, per op, after
 *      Puzzle::CrossSegmentPostProcess, set NextArray                    :64-76
 *   2  the current frontier is expanded with the in-segment hint; children in
 *      this segment set NextArray, and with odd cycles every frontier id
 *      sets CurArray                                                     :78-92
 *   3  without marks and without a frontier the segment has no data and the
 *      step ends with count 0                                            :97
 *   4  the old frontier clears NextArray                                 :101-111
 *   5  ScanBitsAndClear[WithExcl CurArray] in index order: every id is
 *      counted, written to the new frontier (CompressedSegmentWriter,
 *      Common/CompressedFrontier.cpp:128-158: FrontierCompression::Encode of
 *      chunks of StreamVInt::MAX_INDEXES_COUNT values) and expanded with the
 *      cross-segment hint                                                :113-141
 *   6  children in other segments go to SmartMultiplexor::Add(op, seg, idx)
 *      (Common/Multiplexor.cpp:125-161, 200-212): per (op, seg) in emission
 *      order, flushed every 2^12 values through
 *      FrontierCompression::EncodeWithCheck (Common/FrontierCompression.cpp:104-115).
 *
 * The bit arrays are replaced by sorted id lists: marks are sorted and made
 * unique, the old and (odd cycles) current ids are removed by merging, which
 * leaves exactly the set bits ScanBitsAndClearWithExcl visits, in index order.
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define S3_BUCKET_CHUNK 4096u         /* SmartMultiplexor small sector: 2^12 values */
#define S3_FRONTIER_CHUNK (1u << 18)  /* StreamVInt::MAX_INDEXES_COUNT */

static int cmp_u32(const void* a, const void* b) {
    const uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    return x < y ? -1 : x > y ? 1 : 0;
}

typedef struct {
    uint64_t key; /* op * nseg + seg */
    uint64_t ord; /* emission position */
    uint32_t idx;
} s3mark;

static int cmp_s3mark(const void* a, const void* b) {
    const s3mark* x = (const s3mark*)a;
    const s3mark* y = (const s3mark*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->ord < y->ord ? -1 : x->ord > y->ord ? 1 : 0;
}

/* The expansion a plugin performs for ExpandHint {segmentBits, crossSegment}:
 * PancakeOptimizedGPU switches to the in/cross-segment kernels with 29-bit
 * segments and more than 12 pancakes (Puzzles/PancakeOptimizedGpu.cpp:57-79);
 * every other puzzle expands fully. */
int or_hint_mode(const or_puzzle* p, int segment_bits, int cross) {
    if (p->kind == OR_PANCAKE && segment_bits == 29 && p->size > 12)
        return cross ? OR_EXPAND_CROSS_SEGMENT : OR_EXPAND_IN_SEGMENT;
    return OR_EXPAND_FULL;
}

/* the worst case of FrontierCompression::Encode for m values: a bitmap is
 * only chosen when it is smaller than StreamVInt's 4 + 17 per 4 values */
static uint64_t fc_worst(uint64_t m) { return 4 + 17 * ((m + 7) / 8 * 2); }

int64_t or_bfs3_segment_step(const or_puzzle* p, int segment_bits, uint32_t segment, uint32_t nseg,
                             const uint32_t* cur, uint64_t ncur, const uint32_t* old, uint64_t nold,
                             const uint32_t* xseg_idx, const uint64_t* xseg_op_starts, or_seg3_io* io) {
    const int ops = or_operators_count(p);
    const int bf = or_branching_factor(p);
    if (ops <= 0 || segment_bits < 1 || segment_bits > 32 || segment >= nseg) return -1;
    const uint64_t base = (uint64_t)segment << segment_bits;
    const uint64_t in_mask = (1ull << segment_bits) - 1ull;
    const uint64_t nx = xseg_op_starts ? xseg_op_starts[ops] : 0;
    const int odd = or_has_odd_cycles(p);
    const int post = p->kind == OR_PANCAKE && (p->flags & OR_FLAG_INV_INDEX) && segment_bits == 29 && p->size > 12;
    const int mode_in = or_hint_mode(p, segment_bits, 0), mode_x = or_hint_mode(p, segment_bits, 1);
    io->children_in = io->children_cross = io->count = 0;
    io->fr_n = io->fr_nbytes = io->bk_n = io->xs_nbytes = 0;
    if (nx == 0 && ncur == 0) return 0; /* no data (:97) */

    uint32_t* marks = (uint32_t*)malloc((ncur * (uint64_t)bf + nx + 1) * sizeof(uint32_t));
    uint64_t ch[16];
    int32_t chop[16];
    if (!marks) return -1;
    uint64_t nm = 0;
    /* 1 incoming cross-segment marks */
    for (int op = 0; op < ops && nx; op++) {
        const uint64_t b = xseg_op_starts[op], e = xseg_op_starts[op + 1];
        memcpy(marks + nm, xseg_idx + b, (e - b) * sizeof(uint32_t));
        if (post) or_pancake_xseg_postprocess(p->size, (int)segment, marks + nm, e - b);
        nm += e - b;
    }
    /* 2 in-segment children of the current frontier */
    for (uint64_t i = 0; i < ncur; i++) {
        const int k = or_expand_one(p, mode_in, base | cur[i], 0, ch, chop);
        for (int j = 0; j < k; j++) {
            if ((ch[j] >> segment_bits) != segment) continue;
            marks[nm++] = (uint32_t)(ch[j] & in_mask);
            io->children_in++;
        }
    }
    qsort(marks, nm, sizeof(uint32_t), cmp_u32);
    /* 4, 5: unique, minus old, minus cur (odd cycles); cur and old ascending */
    uint64_t fr = 0, io_ = 0, ic = 0;
    for (uint64_t i = 0; i < nm; i++) {
        const uint32_t v = marks[i];
        if (i && marks[i - 1] == v) continue;
        while (io_ < nold && old[io_] < v) io_++;
        if (io_ < nold && old[io_] == v) continue;
        if (odd) {
            while (ic < ncur && cur[ic] < v) ic++;
            if (ic < ncur && cur[ic] == v) continue;
        }
        if (fr >= io->fr_cap) {
            free(marks);
            return -1;
        }
        io->fr_idx[fr++] = v;
    }
    free(marks);
    io->count = fr;
    io->fr_n = fr;

    /* the new frontier: FrontierCompression::Encode per 2^18 values */
    uint64_t pos = 0;
    uint32_t* tv = (uint32_t*)malloc((S3_FRONTIER_CHUNK + 8) * sizeof(uint32_t));
    if (!tv) return -1;
    for (uint64_t c = 0; c < fr; c += S3_FRONTIER_CHUNK) {
        const uint64_t m = (fr - c < S3_FRONTIER_CHUNK) ? fr - c : S3_FRONTIER_CHUNK;
        memcpy(tv, io->fr_idx + c, m * 4);
        if (pos + fc_worst(m) > io->fr_bytes_cap) {
            free(tv);
            return -1;
        }
        pos += (uint64_t)or_fc_encode((int)m, tv, io->fr_bytes + pos);
    }
    free(tv);
    io->fr_nbytes = pos;

    /* 5, 6: cross-segment children of the new frontier, bucketed */
    s3mark* xs = (s3mark*)malloc((fr * (uint64_t)bf + 1) * sizeof(s3mark));
    if (!xs) return -1;
    uint64_t nxs = 0;
    for (uint64_t i = 0; i < fr; i++) {
        const int k = or_expand_one(p, mode_x, base | io->fr_idx[i], 0, ch, chop);
        for (int j = 0; j < k; j++) {
            const uint64_t s = ch[j] >> segment_bits;
            if (s == segment) continue;
            if (s >= nseg) {
                free(xs);
                return -1;
            }
            xs[nxs].key = (uint64_t)chop[j] * nseg + s;
            xs[nxs].ord = nxs;
            xs[nxs].idx = (uint32_t)(ch[j] & in_mask);
            nxs++;
        }
    }
    io->children_cross = nxs;
    qsort(xs, nxs, sizeof(s3mark), cmp_s3mark);
    uint64_t nb = 0, xpos = 0;
    uint32_t* chunk = (uint32_t*)malloc((S3_BUCKET_CHUNK + 16) * sizeof(uint32_t));
    if (!chunk) {
        free(xs);
        return -1;
    }
    int64_t rc = (int64_t)fr;
    for (uint64_t i = 0; i < nxs && rc >= 0;) {
        uint64_t j = i;
        while (j < nxs && xs[j].key == xs[i].key) j++;
        if (nb >= io->bk_cap) {
            rc = -1;
            break;
        }
        io->bk_op[nb] = (uint32_t)(xs[i].key / nseg);
        io->bk_seg[nb] = (uint32_t)(xs[i].key % nseg);
        io->bk_values[nb] = j - i;
        uint64_t bytes = 0;
        for (uint64_t c = i; c < j; c += S3_BUCKET_CHUNK) {
            const uint64_t m = (j - c < S3_BUCKET_CHUNK) ? j - c : S3_BUCKET_CHUNK;
            for (uint64_t t = 0; t < m; t++) chunk[t] = xs[c + t].idx;
            if (xpos + fc_worst(m) > io->xs_cap) {
                rc = -1;
                break;
            }
            const int w = or_fc_encode_with_check((int)m, chunk, io->xs_bytes + xpos);
            xpos += (uint64_t)w;
            bytes += (uint64_t)w;
        }
        io->bk_nbytes[nb] = bytes;
        nb++;
        i = j;
    }
    free(chunk);
    free(xs);
    io->bk_n = nb;
    io->xs_nbytes = xpos;
    return rc;
}
