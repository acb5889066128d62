/*
 * or_segment.c -- one segment step of the disk-based frontier search, CPU
 * restatement.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * DB_OptFS_Solver<4>::Expand(segment) (Common/DiskBasedOptFrontierSearch.cpp:53-132)
 * with the disk stores replaced by in-memory streams:
 *   1  cross-segment marks of the previous step, per op, after
 *      Puzzle::CrossSegmentPostProcess, are OR-ed into the 4-bit op-flag array
 *      (IndexedArray<4>::Set(idx, op), Common/BitArray.h:283-314)      :63-75
 *   2  the frontier (idx, opBits) is expanded with the in-segment hint; every
 *      child in this segment sets its op bit                             :77-99
 *   3  ScanBitsAndClear in index order: count every nonzero entry; entries with
 *      opBits < OperatorsMask are written to the new frontier
 *      (CompressedFrontierWriter, Common/CompressedFrontier.cpp:54-75: StreamVInt
 *      with bounds in chunks of StreamVInt::MAX_INDEXES_COUNT values) and
 *      expanded with the cross-segment hint                               :116-125
 *   4  children in other segments go to SmartMultiplexor::Add(op, seg, idx)
 *      (Common/Multiplexor.cpp:125-134, 200-212): per (op, seg) in emission
 *      order, flushed every 2^12 values through
 *      FrontierCompression::EncodeWithCheck (Common/FrontierCompression.cpp:104-115).
 *
 * The 4-bit array is replaced by a sort of (idx, op) marks, which visits the
 * same nonzero entries in the same index order with the same OR-ed flags.
 * Emission order inside one (op, seg) bucket is parent order, because each
 * parent yields at most one child per op.
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define FS_BUCKET_CHUNK 4096u          /* SmartMultiplexor small sector: 2^12 values */
#define FS_FRONTIER_CHUNK (1u << 18)   /* StreamVInt::MAX_INDEXES_COUNT */

static int cmp_u64(const void* a, const void* b) {
    const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : x > y ? 1 : 0;
}

typedef struct {
    uint64_t key;  /* op * nseg + seg, used for the bucket grouping */
    uint64_t ord;  /* emission position */
    uint32_t idx;
} xmark;

static int cmp_xmark(const void* a, const void* b) {
    const xmark* x = (const xmark*)a;
    const xmark* y = (const xmark*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->ord < y->ord ? -1 : x->ord > y->ord ? 1 : 0;
}

static int xseg_postprocess_applies(const or_puzzle* p, int segment_bits) {
    /* PancakeOptimizedGPU::CrossSegmentPostProcess (Puzzles/PancakeOptimizedGpu.cpp:21-34) */
    return p->kind == OR_PANCAKE && (p->flags & OR_FLAG_INV_INDEX) && segment_bits == 29 && p->size > 12;
}

int64_t or_fs_segment_step(const or_puzzle* p, int segment_bits, uint32_t segment, uint32_t nseg,
                           const uint32_t* idx, const uint8_t* opbits, uint64_t n,
                           const uint32_t* xseg_idx, const uint64_t* xseg_op_starts, or_fs_io* io) {
    const int ops = or_operators_count(p);
    const int bf = or_branching_factor(p);
    if (ops <= 0 || ops > 4 || segment_bits < 1 || segment_bits > 32) return -1;
    const uint32_t mask = (1u << ops) - 1u;
    const uint64_t base = (uint64_t)segment << segment_bits;
    const uint64_t in_mask = (1ull << segment_bits) - 1ull;
    const uint64_t nx = xseg_op_starts ? xseg_op_starts[ops] : 0;

    /* marks: (idx << 8) | op bit, sorted then OR-ed */
    uint64_t cap = n * (uint64_t)bf + nx + 1;
    uint64_t* marks = (uint64_t*)malloc(cap * sizeof(uint64_t));
    uint64_t* ch = (uint64_t*)malloc(16 * sizeof(uint64_t));
    int32_t* chop = (int32_t*)malloc(16 * sizeof(int32_t));
    if (!marks || !ch || !chop) return -1;
    uint64_t nm = 0;
    io->children_in = 0;
    io->children_cross = 0;

    /* 1 cross-segment marks of the previous step */
    if (xseg_op_starts) {
        const int post = xseg_postprocess_applies(p, segment_bits);
        for (int op = 0; op < ops; op++) {
            const uint64_t b = xseg_op_starts[op], e = xseg_op_starts[op + 1];
            uint32_t* tmp = (uint32_t*)malloc((e - b + 1) * sizeof(uint32_t));
            memcpy(tmp, xseg_idx + b, (e - b) * sizeof(uint32_t));
            if (post) or_pancake_xseg_postprocess(p->size, (int)segment, tmp, e - b);
            for (uint64_t i = 0; i < e - b; i++) marks[nm++] = ((uint64_t)tmp[i] << 8) | (1u << op);
            free(tmp);
        }
    }
    /* 2 in-segment expansion of the frontier */
    for (uint64_t i = 0; i < n; i++) {
        const int k = or_expand_one(p, OR_EXPAND_FULL, base | idx[i], opbits[i], ch, chop);
        for (int j = 0; j < k; j++) {
            if ((ch[j] >> segment_bits) != segment) continue;
            marks[nm++] = ((ch[j] & in_mask) << 8) | (1u << chop[j]);
            io->children_in++;
        }
    }
    qsort(marks, nm, sizeof(uint64_t), cmp_u64);

    /* 3 scan in index order; with odd cycles, entries of the current frontier
     * (FrontierArray, :93,117) are dropped.  idx[] is ascending. */
    const int odd = or_has_odd_cycles(p);
    uint64_t count = 0, fr = 0, cur = 0;
    for (uint64_t i = 0; i < nm;) {
        const uint64_t v = marks[i] >> 8;
        uint32_t bits = 0;
        while (i < nm && (marks[i] >> 8) == v) bits |= (uint32_t)(marks[i++] & 0xFF);
        if (odd) {
            while (cur < n && idx[cur] < v) cur++;
            if (cur < n && idx[cur] == v) continue;
        }
        count++;
        if (bits < mask) {
            if (fr >= io->fr_cap) return -1;
            io->fr_idx[fr] = (uint32_t)v;
            io->fr_ops[fr] = (uint8_t)bits;
            fr++;
        }
    }
    free(marks);
    io->count = count;
    io->fr_n = fr;

    /* new frontier stream: StreamVInt with bounds per 2^18 values */
    uint64_t pos = 0;
    {
        uint32_t* tv = (uint32_t*)malloc((FS_FRONTIER_CHUNK + 8) * sizeof(uint32_t));
        uint8_t* tb = (uint8_t*)malloc(FS_FRONTIER_CHUNK + 8);
        for (uint64_t c = 0; c < fr; c += FS_FRONTIER_CHUNK) {
            const uint64_t m = (fr - c < FS_FRONTIER_CHUNK) ? fr - c : FS_FRONTIER_CHUNK;
            memcpy(tv, io->fr_idx + c, m * 4);
            memcpy(tb, io->fr_ops + c, m);
            const uint64_t worst = 4 + 17 * ((m + 7) / 8 * 2) + (m + 7) / 8 * 4;
            if (pos + worst > io->fr_bytes_cap) return -1;
            pos += (uint64_t)or_svint_encode_with_bounds((int)m, tv, tb, io->fr_bytes + pos);
        }
        free(tv);
        free(tb);
    }
    io->fr_nbytes = pos;

    /* 4 cross-segment expansion of the written entries, bucketed */
    xmark* xs = (xmark*)malloc((fr * (uint64_t)bf + 1) * sizeof(xmark));
    uint64_t nxs = 0;
    for (uint64_t i = 0; i < fr; i++) {
        const int k = or_expand_one(p, OR_EXPAND_FULL, base | io->fr_idx[i], io->fr_ops[i], ch, chop);
        for (int j = 0; j < k; j++) {
            const uint64_t s = ch[j] >> segment_bits;
            if (s == segment) continue;
            if (s >= nseg) return -1;
            xs[nxs].key = (uint64_t)chop[j] * nseg + s;
            xs[nxs].ord = nxs;
            xs[nxs].idx = (uint32_t)(ch[j] & in_mask);
            nxs++;
        }
    }
    io->children_cross = nxs;
    qsort(xs, nxs, sizeof(xmark), cmp_xmark);
    uint64_t nb = 0, xpos = 0;
    uint32_t* chunk = (uint32_t*)malloc((FS_BUCKET_CHUNK + 16) * sizeof(uint32_t));
    for (uint64_t i = 0; i < nxs;) {
        uint64_t j = i;
        while (j < nxs && xs[j].key == xs[i].key) j++;
        if (nb >= io->bk_cap) return -1;
        io->bk_op[nb] = (uint32_t)(xs[i].key / nseg);
        io->bk_seg[nb] = (uint32_t)(xs[i].key % nseg);
        io->bk_values[nb] = j - i;
        uint64_t bytes = 0;
        for (uint64_t c = i; c < j; c += FS_BUCKET_CHUNK) {
            const uint64_t m = (j - c < FS_BUCKET_CHUNK) ? j - c : FS_BUCKET_CHUNK;
            for (uint64_t t = 0; t < m; t++) chunk[t] = xs[c + t].idx;
            /* a bitmap is only chosen when smaller than StreamVInt's worst case */
            if (xpos + 4 + 17 * ((m + 7) / 8 * 2) > io->xs_cap) return -1;
            const int w = or_fc_encode_with_check((int)m, chunk, io->xs_bytes + xpos);
            xpos += (uint64_t)w;
            bytes += (uint64_t)w;
            if (xpos > io->xs_cap) return -1;
        }
        io->bk_nbytes[nb] = bytes;
        nb++;
        i = j;
    }
    free(chunk);
    free(xs);
    free(ch);
    free(chop);
    io->bk_n = nb;
    io->xs_nbytes = xpos;
    return (int64_t)count;
}
