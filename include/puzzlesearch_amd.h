/*
 * puzzlesearch_amd.h -- C ABI of the MI355X (gfx950) state-expansion and
 * bit-array frontier library.
 *
 * This is the boundary the reference's PuzzleGpu plugin crosses: every
 * function takes plain pointers and sizes, returns an int status (0 = PS_OK)
 * and never lets a C++ exception escape.  ps_last_error() returns the message
 * of the most recent failure on the calling thread.  Device pointers are HIP
 * device allocations (ps_alloc); streams are hipStream_t values passed as
 * void* (NULL = the default stream).
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   runtime   Puzzles/GPU.h:8-32, Puzzles/GPU.cu:5-58 (ERR, CreateGPUBuffer,
 *             CreateCudaStream, CopyToGpu, CopyFromGpu)
 *   expand    PuzzleGpu::ExpandGpu (Puzzles/PuzzleGpu.h:30) and the kernels it
 *             launches: PancakeExpand / PancakeExpandInSegment /
 *             PancakeExpandCrossSegment (Puzzles/PancakeGpu.h:5-25),
 *             GpuSlidingTilePuzzleSimpleExpand and
 *             GpuSlidingTilePuzzleOptimizedExpand<w,h> (Puzzles/SlidingPuzzleGpu.h:5-18),
 *             GpuHanoiTowersExpand / GpuHanoiTowersExpandOptimized
 *             (Puzzles/HanoiTowersGPU.h:5-19)
 *   bits      BitArray::AndNot / AndNotAndCount / Clear / ScanBits
 *             (Common/BitArray.h:97-165)
 *   drivers   InMemoryThreeBitBFS / InMemoryTwoBitBFS (Common/InMemoryBFS.h:7-9)
 *             run device-resident; DiskBasedOptThreeBitBFS's per-segment
 *             layer structure (Common/DiskBasedOptThreeBitBFS.cpp:54-144)
 *             as the segmented/batched mode.
 */
#ifndef PUZZLESEARCH_AMD_H
#define PUZZLESEARCH_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define PS_OK 0
#define PS_ERR_INVALID_ARGUMENT 1
#define PS_ERR_HIP 2
#define PS_ERR_OUT_OF_MEMORY 3
#define PS_ERR_NOT_INITIALIZED 4
#define PS_ERR_UNSUPPORTED 5
#define PS_ERR_CAPACITY 6
#define PS_ERR_INDEX_RANGE 7
#define PS_ERR_CANCELLED 8
#define PS_ERR_INTERNAL 9 /* a device-side consistency check failed */

typedef void* ps_stream;

/* ---------------------------------------------------------------- puzzles */
/* Puzzle descriptor.  kind selects the reference class whose device I/O layout
 * is reproduced:
 *   PS_PANCAKE         PancakeOptimizedGPU(size, flags & PS_FLAG_INV_INDEX)
 *   PS_SLIDING         SlidingTilePuzzleGpu(size = width, height)
 *   PS_SLIDING_PACKED  SlidingTilePuzzleOptimized<width = size, height>
 *   PS_HANOI4          FourPegHanoiGPU(size = disks, flags & PS_FLAG_SYMMETRY)
 *   PS_HANOI4_OPT      FourPegHanoiOptimizedGPU(size = disks, flags & PS_FLAG_SYMMETRY) */
typedef struct ps_puzzle {
    int32_t kind;
    int32_t size;
    int32_t height;
    int32_t flags;
} ps_puzzle;

enum { PS_PANCAKE = 1, PS_SLIDING = 2, PS_SLIDING_PACKED = 3, PS_HANOI4 = 4, PS_HANOI4_OPT = 5 };
enum { PS_FLAG_SYMMETRY = 1, PS_FLAG_INV_INDEX = 2 };
/* Pancake n > 12 with 29-bit segments: PancakeOptimizedGPU::Expand picks the
 * in-segment (ops 0..10) or cross-segment (ops 11..n-2) kernel from ExpandHint
 * (Puzzles/PancakeOptimizedGpu.cpp:57-79). */
enum { PS_EXPAND_FULL = 0, PS_EXPAND_IN_SEGMENT = 1, PS_EXPAND_CROSS_SEGMENT = 2 };

typedef struct ps_puzzle_info {
    int32_t operators_count;     /* Puzzle::OperatorsCount() */
    int32_t branching_factor;    /* Puzzle::BranchingFactor() */
    int32_t has_odd_cycles;      /* Puzzle::HasOddLengthCycles() */
    int32_t reserved;
    uint64_t indexes_count;      /* Puzzle::IndexesCount() */
} ps_puzzle_info;

/* Host-only query (no GPU needed).  Mirrors Common/Puzzle.h:22-26. */
int ps_puzzle_query(const ps_puzzle* p, ps_puzzle_info* out);
/* Slots per parent written by ps_expand for this mode, or -1 if invalid. */
int ps_expand_slots(const ps_puzzle* p, int32_t mode);

/* ---------------------------------------------------------------- runtime */
/* Select and initialise a device (Puzzles/GPU.cu has no explicit init; the
 * CUDA runtime initialises lazily -- here it is explicit so errors surface). */
int ps_init(int32_t device);
int ps_shutdown(void);
const char* ps_last_error(void);
int ps_device_count(int32_t* count);
/* Device name and properties of the current device. */
int ps_device_props(char* name, int32_t name_cap, int32_t* cu_count, int32_t* lds_bytes_per_block,
                    uint64_t* hbm_bytes);

int ps_stream_create(ps_stream* out);           /* CreateCudaStream  (Puzzles/GPU.cu:27-31) */
int ps_stream_destroy(ps_stream s);             /* DestroyCudaStream (Puzzles/GPU.cu:33-35) */
int ps_stream_sync(ps_stream s);
int ps_device_sync(void);
int ps_alloc(void** out, uint64_t bytes);       /* CreateGPUBuffer   (Puzzles/GPU.cu:12-22) */
int ps_free(void* p);                           /* DestroyGPUBuffer  (Puzzles/GPU.cu:24-26) */
int ps_host_alloc(void** out, uint64_t bytes);  /* pinned host memory */
int ps_host_free(void* p);
/* CopyToGpu (Puzzles/GPU.cu:37-40) */
int ps_memcpy_h2d_async(void* dst, const void* src, uint64_t bytes, ps_stream s);
/* CopyFromGpu without the trailing synchronize (Puzzles/GPU.cu:42-45) */
int ps_memcpy_d2h_async(void* dst, const void* src, uint64_t bytes, ps_stream s);
int ps_memcpy_d2d_async(void* dst, const void* src, uint64_t bytes, ps_stream s);
/* Fill with a byte value using a kernel (safe inside captured graphs). */
int ps_fill_async(void* dst, uint8_t value, uint64_t bytes, ps_stream s);
/* Device-to-device copy by a kernel (16-byte aligned pointers and size); no
 * reference counterpart: it is the known-size copy that calibrates the HBM
 * traffic counters against which the other kernels are measured. */
int ps_copy_async(void* dst, const void* src, uint64_t bytes, ps_stream s);

/* ----------------------------------------------------------- expansion */
/* PuzzleGpu::ExpandGpu (Puzzles/PuzzleGpu.h:30): d_in[i] = index << ops |
 * usedOpBits; d_out[i*S + k] = child << ops | opLabel, or ~0 for an empty
 * slot, S = ps_expand_slots(p, mode).  Bit-identical to the reference kernels
 * (Puzzles/PancakeGpu.cu:92-177, Puzzles/SlidingPuzzleGpu.cu:25-81,210-270,
 * Puzzles/HanoiTowersGPU.cu:114-144,250-281). */
int ps_expand(const ps_puzzle* p, int32_t mode, const uint64_t* d_in, uint64_t* d_out, uint64_t count,
              ps_stream s);
/* PancakeOptimizedGPU::CrossSegmentPostProcess -> PancakeCrossSegmentPostProcessGPU
 * (Puzzles/PancakeOptimizedGpu.cpp:21-34, Puzzles/PancakeGpu.cu:179-193). */
int ps_pancake_xseg_postprocess(int32_t n, int32_t segment, uint32_t* d_idx, uint64_t count, ps_stream s);

/* ----------------------------------------------------------- bit arrays */
/* Bit i of an array lives in 64-bit word i/64 at bit i%64 (Common/BitArray.h:81-83). */
/* a &= ~excl1 [& ~excl2]; *d_count (u64, device) += popcount(a).  excl2 may be NULL.
 * BitArray::AndNot + AndNotAndCount (Common/BitArray.h:97-125). */
int ps_bits_andnot_count(uint64_t* d_a, const uint64_t* d_excl1, const uint64_t* d_excl2, uint64_t nwords,
                         uint64_t* d_count, ps_stream s);
/* *d_count (device) += popcount of the array (BitArray::BitsCount, Common/BitArray.h:109-115). */
int ps_bits_count(const uint64_t* d_a, uint64_t nwords, uint64_t* d_count, ps_stream s);
/* *d_hash (device) += sum over nonzero words w_j of splitmix64(splitmix64(j + word_offset) ^ w_j):
 * the frontier hash of every parity test; no reference counterpart. */
int ps_bits_hash(const uint64_t* d_a, uint64_t nwords, uint64_t word_offset, uint64_t* d_hash, ps_stream s);
/* BitArray::ScanBits (Common/BitArray.h:137-145) into an index list: writes
 * the set bit indexes (ascending) of [0, nbits) to d_out and their number to
 * *d_count (device u64).  Fails with PS_ERR_CAPACITY if more than cap (then
 * the first cap indexes are written).  Allocates a per-call scratch array of
 * one u64 per 2^18 bits and waits for the stream. */
int ps_bits_scan(const uint64_t* d_a, uint64_t nbits, uint64_t* d_out, uint64_t cap, uint64_t* d_count,
                 ps_stream s);
/* ps_bits_scan that allocates nothing and does not wait.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops).  Its chunk counts live in a
 * static device array of 2^20 + 1 u64 (8 MiB, held by every device the library
 * runs on until the process ends), so nbits <= 2^38 (larger: PS_ERR_INVALID_ARGUMENT).
 * *d_count receives the number of set bits; when it exceeds cap only the first
 * cap indexes are written, and the caller checks *d_count (no PS_ERR_CAPACITY).
 * All calls on one device share that array: calls that overlap on one device
 * race, so keep them on one stream or otherwise ordered. */
int ps_bits_scan_noalloc(const uint64_t* d_a, uint64_t nbits, uint64_t* d_out, uint64_t cap, uint64_t* d_count,
                         ps_stream s);
/* Sets bit idx[i] of an array of nbits bits for every i (BitArray::Set in a
 * loop, Common/BitArray.h:81-83).  An index >= nbits is not written: the
 * number of such indexes is added to *d_rejected (device u64; may be NULL). */
int ps_bits_set_list(uint64_t* d_a, uint64_t nbits, const uint64_t* d_idx, uint64_t count, uint64_t* d_rejected,
                     ps_stream s);

/* ----------------------------------------------------------- codecs */
/* Values are handled in chunks: chunk c is values[starts[c], starts[c+1]),
 * at most 2^18 values (StreamVInt::MAX_INDEXES_COUNT).  Encoded chunks are
 * concatenated; chunk c starts at out + offsets[c] and offsets[nchunks] is
 * the total, written by the encoders (device arrays of nchunks+1 u64).  A
 * chunk with no values encodes to no bytes, as in the reference. */

/* StreamVInt::Encode (Common/StreamVInt.cpp:209-231) or, with d_bounds,
 * Encode(count, indexes, bounds, ...) (Common/StreamVInt.cpp:253-279): u32 count, values delta-coded
 * and padded with zeros to a multiple of 8 in 4-value tuples (descriptor byte +
 * 1..4 little-endian bytes each), then 4-bit bounds two per byte.  The input
 * values are not modified. */
int ps_svint_encode(const uint32_t* d_values, const uint8_t* d_bounds, const uint64_t* d_starts, uint32_t nchunks,
                    uint8_t* d_out, uint64_t out_cap, uint64_t* d_offsets, ps_stream s);
/* StreamVInt::Decode (Common/StreamVInt.cpp:233-251, with bounds :280-299) of nchunks concatenated
 * chunks: chunk c is d_in[d_in_offsets[c] .. d_in_offsets[c+1]) (nchunks+1
 * device offsets, as the encoders write them; an empty chunk has no bytes).
 * Chunk c's values go to d_values + value_starts[c]; value_starts (nchunks+1,
 * device) is written. */
int ps_svint_decode(const uint8_t* d_in, const uint64_t* d_in_offsets, uint32_t nchunks, int32_t with_bounds,
                    uint32_t* d_values, uint8_t* d_bounds, uint64_t values_cap, uint64_t* d_value_starts,
                    ps_stream s);
/* OpBitsCompression<bits>::Encode / Decode, bits in {2, 4, 8}; count % 8 == 0
 * (Common/OpBitsCompression.cpp:55-75). */
int ps_opbits_encode(int32_t bits, const uint8_t* d_ops, uint64_t count, uint8_t* d_out, ps_stream s);
int ps_opbits_decode(int32_t bits, const uint8_t* d_in, uint64_t count, uint8_t* d_ops, ps_stream s);
/* FrontierCompression::EncodeWithCheck (with_check = 1) or Encode per chunk:
 * a bitmap chunk (u32 words | 0x80000000, u32 first, bitmap of v - first - 1)
 * when it is smaller, else StreamVInt (Common/FrontierCompression.cpp:28-115). */
int ps_fc_encode(const uint32_t* d_values, const uint64_t* d_starts, uint32_t nchunks, int32_t with_check,
                 uint8_t* d_out, uint64_t out_cap, uint64_t* d_offsets, ps_stream s);
/* FrontierCompression::Decode per chunk (Common/FrontierCompression.cpp:117-125);
 * chunks delimited by nchunks+1 device offsets as for ps_svint_decode. */
int ps_fc_decode(const uint8_t* d_in, const uint64_t* d_in_offsets, uint32_t nchunks, uint32_t* d_values,
                 uint64_t values_cap, uint64_t* d_value_starts, ps_stream s);
/* Groups cross-segment children by (op, target segment) like
 * SmartMultiplexor::Add (Common/Multiplexor.cpp:200-212), in canonical order:
 * bucket b = op * nseg + segment holds the sorted in-segment indexes
 * d_sorted_idx[starts[b] .. starts[b+1]); d_bucket_starts has nops*nseg+1
 * entries. */
int ps_bucket_by_segment(const uint64_t* d_children, const uint8_t* d_ops, uint64_t count, int32_t segment_bits,
                         uint32_t nseg, int32_t nops, uint32_t* d_sorted_idx, uint64_t* d_bucket_starts,
                         ps_stream s);

/* ------------------------------------------------- disk-segment step */
/* One segment step of the disk-based frontier search with 4-bit op flags,
 * DB_OptFS_Solver<4>::Expand(segment) (Common/DiskBasedOptFrontierSearch.cpp:53-132),
 * for puzzles with 1..4 moves.  The workspace holds the segment's op-flag
 * array (2^segment_bits / 2 bytes) and grows its output buffers as needed;
 * one step at a time per workspace. */
typedef struct ps_fs_segment ps_fs_segment;

typedef struct ps_fs_result {
    uint64_t count;           /* states of the new layer in the segment (Expand's return value) */
    uint64_t frontier_count;  /* entries written to the new frontier (a move left unused) */
    uint64_t frontier_bytes;  /* their StreamVInt-with-bounds stream */
    uint64_t frontier_chunks; /* chunks of 2^18 values */
    uint64_t children_in;     /* in-segment children marked */
    uint64_t children_cross;  /* children bucketed for other segments */
    uint64_t buckets;         /* nonempty (op, segment) buckets */
    uint64_t cross_chunks;    /* FrontierCompression chunks of 4096 values over all buckets */
    uint64_t cross_bytes;
    uint64_t scanned_words;   /* op-flag words in the regions the scan visited */
    uint64_t algorithmic_bytes; /* inputs + outputs + 8 per child */
    double device_ms;         /* first to last kernel of the step */
    double wall_ms;
} ps_fs_result;

int ps_fs_segment_create(const ps_puzzle* p, int32_t segment_bits, ps_stream s, ps_fs_segment** out);
int ps_fs_segment_destroy(ps_fs_segment* w);
int ps_fs_segment_info(const ps_fs_segment* w, uint32_t* segments, int32_t* operators, uint64_t* footprint_bytes);
/* Inputs (device): the segment's frontier d_idx[n] (in-segment ids, ascending)
 * with op bits d_ops[n], as CompressedFrontierReader yields them; optionally the
 * cross-segment marks of the previous step, d_xseg_idx[xseg_op_starts[op] ..
 * xseg_op_starts[op+1]) for each op (xseg_op_starts: host, operators+1
 * entries).  Outputs stay in the workspace until the next step. */
int ps_fs_segment_step(ps_fs_segment* w, uint32_t segment, const uint32_t* d_idx, const uint8_t* d_ops, uint64_t n,
                       const uint32_t* d_xseg_idx, const uint64_t* xseg_op_starts, ps_fs_result* res);
/* The new frontier: device list (ids ascending, op bits) and its stream;
 * chunk_offsets (host, frontier_chunks + 1 entries) receives chunk byte offsets.
 * A device pointer is NULL while the workspace has never held such output. */
int ps_fs_segment_frontier(const ps_fs_segment* w, const uint32_t** d_idx, const uint8_t** d_ops,
                           const uint8_t** d_bytes, uint64_t* chunk_offsets, uint64_t cap);
/* The cross-segment buckets in (op, segment) order: bucket b holds values[b]
 * children whose compressed chunks are d_bytes[byte_offsets[b] .. byte_offsets[b+1])
 * (host arrays of cap entries; byte_offsets needs cap + 1).  *d_bytes is NULL
 * while no step of this workspace has produced a cross-segment child. */
int ps_fs_segment_buckets(const ps_fs_segment* w, uint32_t* op, uint32_t* segment, uint64_t* values,
                          uint64_t* byte_offsets, uint64_t cap, const uint8_t** d_bytes);

/* ------------------------------------ disk-segment step of the three-bit BFS */
/* One segment step of DiskBasedOptThreeBitBFS, DB_Opt3BitBFS_Solver::Expand(segment)
 * (Common/DiskBasedOptThreeBitBFS.cpp:54-144), for any puzzle: the incoming
 * cross-segment marks (after Puzzle::CrossSegmentPostProcess) and the
 * in-segment children of the current frontier set NextArray, the old frontier
 * clears it, and the scan of NextArray (excluding the current frontier for
 * puzzles with odd cycles) gives the new layer of the segment.  Its ids are
 * written ascending and as the FrontierCompression::Encode chunks of 2^18
 * values that CompressedSegmentWriter stores; their cross-segment children go
 * to (op, segment) buckets in emission order, each cut into
 * FrontierCompression::EncodeWithCheck chunks of 4096 values as SmartMultiplexor
 * flushes them.  The expansion follows the plugins' ExpandHint: more than 12
 * pancakes in 29-bit segments use the in-segment and cross-segment flip sets.
 * The workspace holds the segment's bit arrays (2^segment_bits bits each,
 * two for puzzles with odd cycles) and grows its output buffers as needed; one
 * step at a time per workspace. */
typedef struct ps_seg3 ps_seg3;

typedef struct ps_seg3_result {
    uint64_t count;           /* states of the new layer in the segment (Expand's return value) */
    uint64_t frontier_bytes;  /* their FrontierCompression stream */
    uint64_t frontier_chunks; /* chunks of 2^18 values */
    uint64_t children_in;     /* in-segment children marked */
    uint64_t children_cross;  /* children bucketed for other segments */
    uint64_t buckets;         /* nonempty (op, segment) buckets */
    uint64_t cross_chunks;    /* FrontierCompression chunks of 4096 values over all buckets */
    uint64_t cross_bytes;
    uint64_t scanned_words;   /* NextArray words in the regions the scan visited */
    uint64_t algorithmic_bytes; /* inputs + outputs + 8 per child + 16 per scanned word */
    double device_ms;         /* first to last kernel of the step */
    double wall_ms;
} ps_seg3_result;

int ps_seg3_create(const ps_puzzle* p, int32_t segment_bits, ps_stream s, ps_seg3** out);
int ps_seg3_destroy(ps_seg3* w);
int ps_seg3_info(const ps_seg3* w, uint32_t* segments, int32_t* operators, uint64_t* footprint_bytes);
/* Inputs (device): the segment's current frontier d_cur[ncur] and old
 * frontier d_old[nold] (in-segment ids, ascending, as CompressedSegmentReader
 * yields them), and optionally the cross-segment marks for this segment,
 * d_xseg[xseg_op_starts[op] .. xseg_op_starts[op+1]) for each op
 * (xseg_op_starts: host, operators+1 entries).  Outputs stay in the workspace
 * until the next step.  A segment with neither marks nor a current frontier
 * has no data: count 0 and no output (:97). */
int ps_seg3_step(ps_seg3* w, uint32_t segment, const uint32_t* d_cur, uint64_t ncur, const uint32_t* d_old,
                 uint64_t nold, const uint32_t* d_xseg, const uint64_t* xseg_op_starts, ps_seg3_result* res);
/* The new frontier: device ids (ascending) and their stream; chunk_offsets
 * (host, frontier_chunks + 1 entries) receives the chunk byte offsets.  A
 * device pointer is NULL while the workspace has never held such output. */
int ps_seg3_frontier(const ps_seg3* w, const uint32_t** d_idx, const uint8_t** d_bytes, uint64_t* chunk_offsets,
                     uint64_t cap);
/* The cross-segment buckets in (op, segment) order: bucket b holds values[b]
 * children, the in-segment ids d_values[value_offsets[b] ..) of its target
 * segment in emission order, compressed as d_bytes[byte_offsets[b] ..
 * byte_offsets[b+1]) (host arrays of cap entries; the offset arrays need
 * cap + 1).  Device pointers are NULL while no step has produced a child. */
int ps_seg3_buckets(const ps_seg3* w, uint32_t* op, uint32_t* segment, uint64_t* values, uint64_t* value_offsets,
                    uint64_t* byte_offsets, uint64_t cap, const uint32_t** d_values, const uint8_t** d_bytes);

/* ------------------------------------------- HanoiTowers vertical solver */
/* 4-peg Hanoi states reduced by the symmetry of pegs 1..3 (pegs 1..3 ordered
 * by bottom disk, largest first, empty pegs last); an index holds disk d's peg
 * in bits 2d..2d+1. */

/* kernel_expand_inseg / kernel_expand_inseg_without_smallest launched by
 * GpuExpandInSegment<size> / GpuExpandInSegmentWithoutSmallest<size>
 * (HanoiTowers/HanoiExpandGPU.cu:98-229, :231-275): for each id of the 2^32-state
 * segment, d_children[i*6 + k] (or i*3 + k without the smallest disk's moves)
 * is the in-segment id of the child of move k, moves (0,1) (0,2) (0,3) (1,2)
 * (1,3) (2,3) (without: the three pairs of the pegs other than the smallest
 * disk's), or the parent's own id where the child leaves the segment.
 * disks 1..24; the segment must lie in the index space.  Enqueued on s. */
int ps_hanoi_expand_in_segment(int32_t disks, uint32_t segment, const uint32_t* d_idx, uint64_t count,
                               uint32_t* d_children, int32_t without_smallest, ps_stream s);

/* HanoiTowers::Expand(indexes, children) (HanoiTowers/HanoiTowers.cpp:152-172,
 * :322-327): d_children[i*6 + k] is the reduced child of d_idx[i] by move k,
 * moves (0,1) (0,2) (0,3) (1,2) (1,3) (2,3); a move with no disk to move gives
 * the parent itself.  Indexes must lie in the disks' index space. */
int ps_hanoi_expand(int32_t disks, const uint64_t* d_idx, uint64_t count, uint64_t* d_children, ps_stream s);

/* HanoiSearch<size> (HanoiTowers/HanoiSearch.cpp:46-290): the three-bit BFS of
 * the reduced states over segments of 2^segment_bits states (even, 8..32; 0 =
 * 32 as in the reference), with only the segments a layer reaches held in HBM
 * (three bit arrays of 2^segment_bits bits each) instead of disk stores. */
typedef struct ps_hanoi_search ps_hanoi_search;
typedef struct ps_hanoi_search_stats {
    uint64_t layers;
    uint64_t segments_allocated;  /* segments held in HBM */
    uint64_t expanded;            /* reduced parents expanded */
    uint64_t children;            /* children generated (the four-bit shortcut counts 3) */
    uint64_t listed_children;     /* children outside their parent's 2^18-state tile */
    uint64_t algorithmic_bytes;
    uint64_t hbm_bytes;
    double device_ms, wall_ms;
} ps_hanoi_search_stats;
/* hbm_budget 0: three quarters of the free HBM at creation */
int ps_hanoi_search_create(int32_t disks, int32_t segment_bits, uint64_t hbm_budget, ps_stream s,
                           ps_hanoi_search** out);
int ps_hanoi_search_destroy(ps_hanoi_search* h);
/* Runs from `start` (a reduced state) for at most max_layers layers.
 * restored[l] is the reference's result (layer counts restored to the
 * unreduced puzzle with PegsMult, HanoiTowers/HanoiTowers.h:48-55; layer 0 is
 * 1), reduced[l] the reduced states, hashes[l] (may be NULL) the hash of layer
 * l's reduced frontier as a bit array over the whole index space (sum over
 * nonzero words w_j of splitmix64(splitmix64(j) ^ w_j)).  Fails with
 * PS_ERR_CAPACITY when a reached segment does not fit the budget. */
int ps_hanoi_search_run(ps_hanoi_search* h, uint64_t start, int32_t max_layers, uint64_t* restored,
                        uint64_t* reduced, uint64_t* hashes, int32_t cap, int32_t* n_layers);
int ps_hanoi_search_stats_get(const ps_hanoi_search* h, ps_hanoi_search_stats* st);

/* --------------------------------------------------- multi-search engine */
/* Continuous batching of complete three-bit BFS jobs (InMemoryThreeBitBFS,
 * Common/InMemoryThreeBitBFS.cpp:4-48, for many puzzles at once).  No
 * counterpart in the reference: it serves a stream of search requests.  Jobs
 * of the puzzles compiled into the multi-search kernels (sliding tile with 8,
 * 9, 10 or 12 cells, pancake 8..12, 4-peg Hanoi up to 16 disks) are queued by
 * ps_batch_submit.  ps_batch_step admits queued jobs first come first served
 * while their HBM footprint fits the budget, advances the running jobs, and
 * retires finished jobs.  A finished job's device search is kept for the next
 * job of the same puzzle and index space and released only when a queued job
 * needs its memory (releasing device memory waits for the whole device).  Not
 * thread-safe: one caller drives a batch. */
typedef struct ps_batch ps_batch;

enum { PS_JOB_QUEUED = 0, PS_JOB_RUNNING = 1, PS_JOB_DONE = 2, PS_JOB_FAILED = 3 };

/* PS_BATCH_STREAMS: every job runs on its own stream with its own completion
 * event.  A step enqueues the next layers of each job that has none in flight
 * (4 at first, then up to 32, replayed from the job's captured layer graph)
 * and returns once at least one job's layers have completed, so a job is never
 * held back by a larger one and the jobs' kernels overlap.
 * PS_BATCH_FUSED: all running jobs share one persistent expand, one finalize
 * and one bookkeeping launch per step; currently disabled (fails with
 * PS_ERR_UNSUPPORTED). */
enum { PS_BATCH_STREAMS = 0, PS_BATCH_FUSED = 1 };

typedef struct ps_batch_options {
    int32_t max_jobs;      /* jobs running at once (0: 64) */
    int32_t segment_bits;  /* segments of jobs larger than one segment (default 22; 0: none) */
    uint64_t hbm_budget;   /* bytes running jobs may hold (0: half of the free HBM at creation) */
    int32_t mode;          /* PS_BATCH_STREAMS (default) or PS_BATCH_FUSED */
    int32_t job_grid_blocks; /* PS_BATCH_STREAMS: most blocks of a job's persistent kernels, so that a
                              * large job leaves room on the device for the others (0: no cap) */
    int32_t reserved_cus;    /* PS_BATCH_STREAMS: CUs 0..reserved_cus-1 are kept free of jobs of at
                              * least 2^25 states (their streams are CU-masked), so small jobs never
                              * wait for a large job's persistent kernel (0: none) */
    int32_t shortest_first;  /* PS_BATCH_STREAMS: a job whose unvisited states exceed this many times
                              * those of the smallest running job gets no new layers, for at most
                              * 20 ms at a time (default 4; 0: every job advances) */
} ps_batch_options;

typedef struct ps_batch_job_info {
    int32_t state;      /* PS_JOB_* */
    int32_t layers;     /* layer counts available (state DONE) */
    uint64_t footprint; /* HBM bytes while running */
    uint64_t steps;     /* engine steps the job ran in */
    uint64_t expanded;  /* parents expanded */
    uint64_t children;
    uint64_t submit_ns, admit_ns, done_ns; /* steady-clock times */
} ps_batch_job_info;

typedef struct ps_batch_stats {
    uint64_t steps;
    uint64_t jobs_done;
    uint64_t states_done;  /* sum of the layer counts of finished jobs */
    uint64_t expanded;
    uint64_t hbm_used, hbm_budget; /* used: running jobs and the finished jobs' searches kept for reuse */
    double step_ms;        /* host time spent in steps that ran jobs */
} ps_batch_stats;

void ps_batch_options_default(ps_batch_options* o);
int ps_batch_create(const ps_batch_options* opt, ps_stream s, ps_batch** out);
int ps_batch_destroy(ps_batch* b);
/* index_count 0: the puzzle's IndexesCount(); fails if the footprint exceeds the budget */
int ps_batch_submit(ps_batch* b, const ps_puzzle* p, uint64_t start, uint64_t index_count, uint64_t* job_id);
int ps_batch_step(ps_batch* b, int32_t* running, int32_t* finished, int32_t* queued);
/* ids of the jobs that finished (or failed) in the last step */
int ps_batch_finished(const ps_batch* b, uint64_t* ids, int32_t cap, int32_t* n);
/* a job's state and, once done, its layer counts and frontier hashes (cap entries each) */
int ps_batch_job(const ps_batch* b, uint64_t job_id, ps_batch_job_info* info, uint64_t* counts, uint64_t* hashes,
                 int32_t cap);
/* drop a finished job's results */
int ps_batch_forget(ps_batch* b, uint64_t job_id);
int ps_batch_stats_get(const ps_batch* b, ps_batch_stats* out);

/* ----------------------------------------------------------- device BFS */
typedef struct ps_bfs ps_bfs;

enum { PS_BFS_BATCHED = 0, PS_BFS_PLAIN = 1 };
/* How the three-bit expand kernel marks a child outside its parent's LDS
 * window: PS_CROSS_BYTES stores 1 into a byte-per-state array that the
 * finalize pass folds into the next layer (no read-modify-write traffic; costs
 * one byte of HBM per state), PS_CROSS_ATOMIC ORs the bit into the next-layer
 * array with a device-scope atomic.  PS_CROSS_LISTS (10..12 pancakes, batched
 * or plain mode, top-down layers, the default 9!-rank tiles, index_count >=
 * n!) appends the child to a list owned by the parent's tile and the child's
 * tile; a per-tile pass ORs every list into the next layer.  Each of the n!/9!
 * tiles owns (n-9)*9! four-byte entries: 5.75 GB for 12 pancakes. */
enum { PS_CROSS_BYTES = 0, PS_CROSS_ATOMIC = 1, PS_CROSS_LISTS = 2 };
/* Layer direction of the three-bit search.  Top-down: every frontier state
 * marks its children.  Bottom-up: every state not yet visited looks for a
 * parent in the frontier among its neighbours (all moves here are
 * reversible), which makes no scattered writes.  Auto goes bottom-up for a
 * layer whose frontier outnumbers the states not yet visited.  Bottom-up
 * layers need every index to be a state: ps_bfs_create refuses
 * PS_DIR_BOTTOM_UP (PS_ERR_UNSUPPORTED) for symmetry-reduced Hanoi, the packed
 * sliding index and more than 12 pancakes, where PS_DIR_AUTO stays top-down. */
enum { PS_DIR_TOP_DOWN = 0, PS_DIR_AUTO = 1, PS_DIR_BOTTOM_UP = 2 };

typedef struct ps_bfs_options {
    int32_t bits;           /* 3: InMemoryThreeBitBFS semantics, 2: InMemoryTwoBitBFS */
    int32_t segment_bits;   /* 0: one segment; else segments of 2^segment_bits states */
    int32_t mode;           /* PS_BFS_BATCHED or PS_BFS_PLAIN */
    int32_t max_layers;     /* 0: until a layer is empty */
    uint64_t index_count;   /* 0: Puzzle::IndexesCount() */
    int32_t tile_bits_log2; /* 0: tuned default; else LDS window of 2^k states (-1: no window) */
    int32_t block_threads;  /* 0: tuned default (1024 for batched PS_CROSS_LISTS, else 512);
                               256 or 512 (1024 with PS_CROSS_LISTS) */
    int32_t grid_blocks;    /* 0: tuned default (persistent grid) */
    int32_t hash;           /* 1: per-layer frontier hashes */
    int32_t graph;          /* 1: capture the batched layer loop in a hipGraph (default on) */
    int32_t graph_layers;   /* layers per captured graph (0: default 4) */
    int32_t cross_marks;    /* PS_CROSS_BYTES (default), PS_CROSS_ATOMIC or PS_CROSS_LISTS, see above */
    int32_t skip_visited;   /* 1: a child outside the LDS window whose bit is set in the
                               current or previous layer is not marked (it cannot be new:
                               in an undirected graph children of layer L lie in L-1..L+1) */
    int32_t direction;      /* PS_DIR_TOP_DOWN (default), PS_DIR_AUTO or PS_DIR_BOTTOM_UP;
                               batched mode with an LDS window only */
    int32_t fused_layers;   /* 1: one launch per layer finalizes and expands each tile
                               (k_bfs3_tile_step, four bit arrays instead of three);
                               batched three-bit PS_CROSS_ATOMIC top-down searches without
                               skip_visited whose tiles lie inside one segment */
} ps_bfs_options;

typedef struct ps_bfs_stats {
    uint64_t layers;
    uint64_t expanded;            /* parents expanded, all layers */
    uint64_t children;            /* children generated (marks), all layers */
    uint64_t algorithmic_bytes;   /* see the header of bfs.hip */
    uint64_t expand_bytes;        /* algorithmic bytes of the expand kernels alone */
    uint64_t scan_words;          /* frontier words read by the expand kernels */
    uint64_t expand_launches;
    uint64_t finalize_launches;
    double expand_ms;             /* summed event time of the expand kernels */
    double finalize_ms;
    double total_ms;              /* event time of the whole run */
    uint64_t pull_words;          /* words scanned by bottom-up layers (old and cur read) */
    uint64_t pull_probes;         /* frontier bits read outside the tile by bottom-up layers */
    double apply_ms;              /* summed event time of the cross-list apply kernels */
    uint64_t apply_launches;      /* batched PS_CROSS_LISTS runs one fused kernel per layer, counted
                                     in expand_launches (apply and finalize launches stay 0) */
    uint64_t list_entries;        /* cross-tile list entries written (PS_CROSS_LISTS), all layers */
    double pull_ms;               /* summed event time of the bottom-up kernel of batched PS_CROSS_LISTS
                                     runs (k_pkl_pull; its launches are not in expand_launches) */
    uint64_t pull_launches;
    uint64_t pull_bytes;          /* algorithmic bytes of those launches (expand_bytes excludes them) */
} ps_bfs_stats;

void ps_bfs_options_default(ps_bfs_options* o);
int ps_bfs_create(const ps_puzzle* p, const ps_bfs_options* opt, ps_stream s, ps_bfs** out);
int ps_bfs_destroy(ps_bfs* b);
/* Runs the search from start_index to exhaustion (or max_layers).  counts[l]
 * and hashes[l] (may be NULL) receive layer l; *n_layers the layer count. */
int ps_bfs_run(ps_bfs* b, uint64_t start_index, uint64_t* counts, uint64_t* hashes, int32_t cap,
               int32_t* n_layers);
/* Stepwise use: reset to a start state, then one layer per call (*count = 0
 * when the search is exhausted). */
int ps_bfs_reset(ps_bfs* b, uint64_t start_index);
int ps_bfs_step(ps_bfs* b, uint64_t* count, uint64_t* hash);
int ps_bfs_stats_get(const ps_bfs* b, ps_bfs_stats* out);
/* Device bit array of the current frontier (three-bit) or the 2-bit array. */
int ps_bfs_frontier(const ps_bfs* b, const uint64_t** d_words, uint64_t* nwords);
/* HBM bytes ps_bfs_create would allocate for this puzzle and options. */
int ps_bfs_footprint(const ps_puzzle* p, const ps_bfs_options* opt, uint64_t* bytes);

#ifdef __cplusplus
}
#endif
#endif
