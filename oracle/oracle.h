/*
 * oracle.h -- CPU restatement of lightln2/PuzzleSearch's state-expansion and
 * bit-array frontier path, in plain C.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code.  It may be used only by tests/, by __graft_entry__.smoke()
 * and by bench.py's cpu_baseline leg, always as the checker (or the timed CPU
 * baseline), never as the thing measured on the GPU.
 *
 * Every function names the reference file:line it restates.  How each part is
 * pinned (golden vectors, known-answer tests, closed forms, brute force) is
 * listed in the header comment of the matching test file under tests/.
 */
#ifndef PS_ORACLE_H
#define PS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_INVALID_INDEX UINT64_MAX

/* Puzzle kinds; the numbering matches ps_puzzle in include/puzzlesearch_amd.h */
enum {
    OR_PANCAKE = 1,          /* PancakeOptimized(GPU): size = n, flags&2 = inverse index */
    OR_SLIDING = 2,          /* SlidingTilePuzzleGpu / ...Simple: size = width, height = height */
    OR_SLIDING_PACKED = 3,   /* SlidingTilePuzzleOptimized<width,height> */
    OR_HANOI4 = 4,           /* FourPegHanoiGPU (linked-list kernel): size = disks, flags&1 = symmetry */
    OR_HANOI4_OPT = 5,       /* FourPegHanoiOptimizedGPU (bit-plane kernel) */
};
enum { OR_FLAG_SYMMETRY = 1, OR_FLAG_INV_INDEX = 2 };
/* Expansion modes of the pancake 29-bit segmented rank (n > 12 only). */
enum { OR_EXPAND_FULL = 0, OR_EXPAND_IN_SEGMENT = 1, OR_EXPAND_CROSS_SEGMENT = 2 };

typedef struct or_puzzle {
    int32_t kind;
    int32_t size;
    int32_t height;
    int32_t flags;
} or_puzzle;

/* ---- hashing (canonical forms used by every parity test) ---- */
uint64_t or_splitmix64(uint64_t x);
/* Order-free hash of a bit array in BitArray word layout (bit i in u64 word
 * i/64 at bit i%64): sum over nonzero words w_j of splitmix64(splitmix64(j) ^ w_j). */
uint64_t or_bits_hash(const uint64_t* words, uint64_t nwords);
/* Order-free multiset hash of expanded (child, op) pairs:
 * sum of splitmix64((child << 4) ^ op) over non-INVALID children. */
uint64_t or_children_hash(const uint64_t* children, const int32_t* ops, uint64_t count);

/* ---- permutation rank / unrank (Puzzles/PermutationMap.cpp:7-58) ---- */
void or_perm_compact(int* arr, int size);
void or_perm_uncompact(int* arr, int size);
uint64_t or_perm_rank(int* arr, int size);           /* compacts arr in place */
void or_perm_unrank(uint64_t index, int* arr, int size);

/* ---- pancake, 29-bit segmented rank (Puzzles/PancakeOptimized.cpp:26-77) ---- */
uint64_t or_pancake_rank(int* arr, int n);
void or_pancake_unrank(uint64_t index, int* arr, int n);

/* ---- sliding tile packed encoding (Puzzles/SlidingTilePuzzleOptimized.cpp:79-192) ---- */
/* Standard state (tile per cell, 0 = blank) -> packed (segment << 32 | idx); -1 on bad parity. */
int64_t or_packed_from_standard(int width, int height, const int* cells);
void or_packed_to_standard(int width, int height, uint64_t index, int* cells);

/* ---- generic puzzle interface (Common/Puzzle.h:16-48) ---- */
int or_operators_count(const or_puzzle* p);
int or_branching_factor(const or_puzzle* p);
uint64_t or_indexes_count(const or_puzzle* p);
int or_has_odd_cycles(const or_puzzle* p);
/* Parse / ToString in the reference's text formats. Parse returns OR_INVALID_INDEX on error. */
uint64_t or_parse(const or_puzzle* p, const char* text);
int or_to_string(const or_puzzle* p, uint64_t index, char* buf, int cap);

/* Device I/O layout of PuzzleGpu::ExpandGpu (Puzzles/PuzzleGpu.cpp:53-82 and the
 * kernels it launches): in[i] = index << OperatorsCount() | usedOpBits,
 * out[i*S + k] = child << OperatorsCount() | opLabel, or ~0.  Returns S (slots per
 * parent) or -1 if the puzzle/mode pair is not valid. */
int or_expand_slots(const or_puzzle* p, int mode, const uint64_t* in, uint64_t count, uint64_t* out);
int or_slots_per_parent(const or_puzzle* p, int mode);
/* Reference CPU-twin Expand of one parent: children in emission order (slot
 * order with INVALID slots dropped).  Returns the number of children. */
int or_expand_one(const or_puzzle* p, int mode, uint64_t index, int op_bits,
                  uint64_t* children, int32_t* ops);
/* FourPegHanoiOptimized::Expand (CPU class): restores symmetry only when a peg
 * became empty or non-empty (Puzzles/FourPegHanoiOptimized.cpp:207-238). */
int or_hanoi_expand_cpu_opt(int disks, int symmetry, uint64_t index, int op_bits,
                            uint64_t* children, int32_t* ops);
/* PuzzleGpu::CombineIndexAndOpBits / SplitIndexAndOps (Puzzles/PuzzleGpu.cpp:35-51). */
void or_combine_index_opbits(int ops, const uint64_t* idx, const int32_t* op_bits, uint64_t count, uint64_t* out);
void or_split_index_ops(int ops, const uint64_t* packed, uint64_t count, uint64_t* idx, int32_t* op);
/* PancakeOptimized::CrossSegmentPostProcess, GPU form (Puzzles/PancakeGpu.cu:179-193). */
void or_pancake_xseg_postprocess(int n, int segment, uint32_t* idx, uint64_t count);

/* ---- HanoiTowers vertical solver (HanoiTowers/) ----
 * States are 4-peg Hanoi indexes (disk d's peg in bits 2d..2d+1) reduced by
 * the symmetry of pegs 1..3: pegs 1..3 ordered by bottom disk, largest first. */
/* HanoiTowers::PegsMult (HanoiTowers/HanoiTowers.h:48-55): the unreduced states a
 * reduced one stands for. */
int or_hv_pegs_mult(uint64_t index);
/* HanoiTowers::InSegmentOnly (:37-46) over the low seg_bits bits (32 in the
 * reference): no disk above the segment's disks can move. */
int or_hv_in_segment_only(uint64_t index, int seg_bits);
/* HanoiTowers::Expand(index) (HanoiTowers/HanoiTowers.cpp:152-172): the six
 * moves (0,1) (0,2) (0,3) (1,2) (1,3) (2,3), symmetry restored where a peg is
 * or becomes empty. */
void or_hv_expand(int disks, uint64_t index, uint64_t* out6);
/* HanoiTowers::ExpandCrossSegment (:174-198): the children outside the segment
 * of 2^seg_bits states, in move order; returns their number. */
int or_hv_expand_cross_segment(int disks, int seg_bits, uint64_t segment, uint32_t index, uint64_t* out);
/* kernel_expand_inseg / kernel_expand_inseg_without_smallest
 * (HanoiTowers/HanoiExpandGPU.cu:98-229), 2^32-state segments: 6 (or 3) slots
 * per parent, the parent's own id where the child leaves the segment. */
void or_hv_expand_in_segment_slots(int disks, uint32_t segment, const uint32_t* idx, uint64_t count,
                                   int without_smallest, uint32_t* out);
/* HanoiSearch (HanoiTowers/HanoiSearch.cpp:46-290) with segments of 2^seg_bits
 * states (even, 6..32; 32 in the reference) and in-memory stores.  restored[l]
 * is the reference's result (PegsMult-restored counts, layer 0 = 1), reduced[l]
 * the reduced states, hashes[l] (may be NULL) the hash of layer l's reduced
 * frontier as a bit array over the whole index space.  Returns the number of
 * layers, at most max_layers, or -1. */
int or_hv_search(int disks, uint64_t start, int seg_bits, int max_layers, uint64_t* restored, uint64_t* reduced,
                 uint64_t* hashes);

/* ---- bit arrays (Common/BitArray.h:69-169) ---- */
uint64_t or_bits_andnot_count(uint64_t* a, const uint64_t* excl, uint64_t nwords);
void or_bits_andnot(uint64_t* a, const uint64_t* excl, uint64_t nwords);
uint64_t or_bits_count(const uint64_t* a, uint64_t nwords);

/* ---- BFS drivers ----
 * Per-layer counts (result vector of the reference drivers) and per-layer
 * frontier hashes.  Returns the number of layers (>=1) or -1 on error.
 * index_count == 0 uses or_indexes_count(p). */
/* InMemoryThreeBitBFS (Common/InMemoryThreeBitBFS.cpp:4-48). hashes[l] = hash of
 * layer l's frontier bit array. */
int or_bfs3(const or_puzzle* p, uint64_t start, uint64_t index_count, int max_layers,
            uint64_t* counts, uint64_t* hashes);
/* InMemoryTwoBitBFS (Common/InMemoryTwoBitBFS.cpp:24-69).  hashes[l] = hash of the
 * whole 2-bit array after layer l (value v of state i at bits 2i..2i+1, natural
 * order: bit 2i = v & 1). */
int or_bfs2(const or_puzzle* p, uint64_t start, uint64_t index_count, int max_layers,
            uint64_t* counts, uint64_t* hashes);
/* DiskBasedOptThreeBitBFS with the disk stores replaced by in-memory lists
 * (Common/DiskBasedOptThreeBitBFS.cpp:54-144, 164-252). */
int or_bfs3_segmented(const or_puzzle* p, uint64_t start, int segment_bits, int max_layers,
                      uint64_t* counts);

/* One three-bit layer step for timing: expands every parent set in `cur` whose
 * index is in [first, first + span) and marks children in `next`.  Returns the
 * number of parents expanded; *children_out receives the number of marks. */
uint64_t or_bfs3_expand_range(const or_puzzle* p, const uint64_t* cur, uint64_t* next,
                              uint64_t first, uint64_t span, uint64_t* children_out);

/* ---- StreamVInt (Common/StreamVInt.cpp:149-300) ---- */
int or_svint_encode_tuple(const uint32_t* v4, uint8_t* buf);
int or_svint_decode_tuple(const uint8_t* buf, uint32_t* v4);
void or_svint_encode_bounds_tuple(const uint8_t* b8, uint8_t* buf);
void or_svint_decode_bounds_tuple(const uint8_t* buf, uint8_t* b8);
int or_svint_encode_indexes(int count, const uint32_t* idx, uint8_t* buf);
int or_svint_decode_indexes(int count, const uint8_t* buf, uint32_t* idx);
int or_svint_decode_indexes_diff(int count, const uint8_t* buf, uint32_t* idx);
/* Encode mutates idx to deltas and pads it to a multiple of 8 (idx needs that room). */
int or_svint_encode(int count, uint32_t* idx, uint8_t* buf);
int or_svint_decode(int* size, const uint8_t* buf, uint32_t* idx);
int or_svint_encode_with_bounds(int count, uint32_t* idx, uint8_t* bounds, uint8_t* buf);
int or_svint_decode_with_bounds(int* size, const uint8_t* buf, uint32_t* idx, uint8_t* bounds);

/* ---- OpBitsCompression<2/4/8> (Common/OpBitsCompression.cpp:10-76) ---- */
int64_t or_opbits_encode(int bits, uint64_t count, const uint8_t* ops, uint8_t* buf);
int64_t or_opbits_decode(int bits, uint64_t count, const uint8_t* buf, uint8_t* ops);

/* ---- FrontierCompression (Common/FrontierCompression.cpp:28-150) ---- */
int or_fc_is_bitmap(const uint8_t* buf);
int64_t or_fc_bitmap_size(int count, const uint32_t* idx);
int or_fc_encode_bitmap(int count, const uint32_t* idx, uint8_t* buf);
int or_fc_encode_bitmap_with_check(int count, const uint32_t* idx, uint8_t* buf);
int or_fc_decode_bitmap(int* size, const uint8_t* buf, uint32_t* idx);
int or_fc_encode(int count, uint32_t* idx, uint8_t* buf);
int or_fc_encode_with_check(int count, uint32_t* idx, uint8_t* buf);
int or_fc_decode(int* size, const uint8_t* buf, uint32_t* idx);

/* ---- one segment step of the disk-based frontier search ----
 * DB_OptFS_Solver<4>::Expand(segment) (Common/DiskBasedOptFrontierSearch.cpp:53-132)
 * with in-memory streams.  Inputs: the segment's frontier (idx ascending,
 * opBits), and optionally the cross-segment marks of the previous step
 * (xseg_idx[xseg_op_starts[op] .. xseg_op_starts[op+1]) per op).  Outputs
 * (caller buffers): the new frontier list and its StreamVInt-with-bounds
 * stream (2^18-value chunks); the cross-segment buckets in (op, seg) order,
 * each as its FrontierCompression::EncodeWithCheck chunks of 4096 values in
 * emission order, concatenated in xs_bytes.  Returns the layer count of the
 * segment or -1 (capacity or argument error). */
typedef struct or_fs_io {
    uint32_t* fr_idx;
    uint8_t* fr_ops;
    uint64_t fr_cap;
    uint64_t fr_n;
    uint8_t* fr_bytes;
    uint64_t fr_bytes_cap;
    uint64_t fr_nbytes;
    uint32_t* bk_op;
    uint32_t* bk_seg;
    uint64_t* bk_values;
    uint64_t* bk_nbytes;
    uint64_t bk_cap;
    uint64_t bk_n;
    uint8_t* xs_bytes;
    uint64_t xs_cap;
    uint64_t xs_nbytes;
    uint64_t children_in;
    uint64_t children_cross;
    uint64_t count;
} or_fs_io;

int64_t or_fs_segment_step(const or_puzzle* p, int segment_bits, uint32_t segment, uint32_t nseg,
                           const uint32_t* idx, const uint8_t* opbits, uint64_t n,
                           const uint32_t* xseg_idx, const uint64_t* xseg_op_starts, or_fs_io* io);

/* ---- one segment step of the disk-based three-bit BFS ----
 * DB_Opt3BitBFS_Solver::Expand(segment) (Common/DiskBasedOptThreeBitBFS.cpp:54-144)
 * with in-memory lists.  Inputs: the segment's current and old frontiers (ids
 * ascending) and the cross-segment marks of the previous step per op
 * (xseg_idx[xseg_op_starts[op] .. xseg_op_starts[op+1])).  Outputs: the new
 * frontier (ids ascending) and its FrontierCompression::Encode chunks of 2^18
 * values (CompressedSegmentWriter); the cross-segment buckets in (op, seg)
 * order, each as FrontierCompression::EncodeWithCheck chunks of 4096 values
 * in emission order, concatenated in xs_bytes.  Returns the layer count of the
 * segment or -1 (capacity or argument error). */
typedef struct or_seg3_io {
    uint32_t* fr_idx;
    uint64_t fr_cap;
    uint64_t fr_n;
    uint8_t* fr_bytes;
    uint64_t fr_bytes_cap;
    uint64_t fr_nbytes;
    uint32_t* bk_op;
    uint32_t* bk_seg;
    uint64_t* bk_values;
    uint64_t* bk_nbytes;
    uint64_t bk_cap;
    uint64_t bk_n;
    uint8_t* xs_bytes;
    uint64_t xs_cap;
    uint64_t xs_nbytes;
    uint64_t children_in;
    uint64_t children_cross;
    uint64_t count;
} or_seg3_io;

int64_t or_bfs3_segment_step(const or_puzzle* p, int segment_bits, uint32_t segment, uint32_t nseg,
                             const uint32_t* cur, uint64_t ncur, const uint32_t* old, uint64_t nold,
                             const uint32_t* xseg_idx, const uint64_t* xseg_op_starts, or_seg3_io* io);
/* The expansion mode of ExpandHint {segment_bits, cross} for this puzzle
 * (Puzzles/PancakeOptimizedGpu.cpp:57-79). */
int or_hint_mode(const or_puzzle* p, int segment_bits, int cross);

#ifdef __cplusplus
}
#endif
#endif
