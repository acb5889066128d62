// One disk-segment step of the frontier search on the device:
// DB_OptFS_Solver<4>::Expand(segment) (Common/DiskBasedOptFrontierSearch.cpp:53-132)
// with the segment's frontier and incoming cross-segment marks already in HBM.
//
//   mark     in-segment children of the frontier and the incoming marks, as
//            the marks of the 4-bit op-flag array                    (:63-99):
//            counted per region of 2^18 states, placed in a mark list by
//            region (fs_kernels.h)
//   scan     ScanBitsAndClear (:116-123) per region in LDS: the region's
//            flags built from its marks and counted, an exclusive scan of
//            the output counts, then the flags rebuilt and the (id, op bits)
//            of the entries with an unused move written in id order
//   stream   the new frontier as StreamVInt with bounds in 2^18-value chunks,
//            as CompressedFrontierWriter writes it (Common/CompressedFrontier.cpp:54-75)
//   cross    (a second stream, beside `stream`) children of the written
//            entries in other segments, compacted in emission order, stably
//            grouped by (op, segment) with a counting sort (seg_sort.h), and
//            each bucket cut into 4096-value chunks compressed with
//            FrontierCompression::EncodeWithCheck, as SmartMultiplexor flushes
//            them (Common/Multiplexor.cpp:125-161)
//
// The host waits for the region counts, the cross-child count and the bucket
// list; the chunk offsets of both streams are read back with the step's end.
//
// Emission order inside a bucket is (parent, slot) order, the order in which
// ExpandBuffer hands children to SmartMultiplexor::Add for one segment, so the
// bucket bytes equal the reference's for a segment step run on its own.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "codec_internal.h"
#include "devbuf.h"
#include "fs_launch.h"
#include "ps_internal.h"
#include "seg_sort.h"

namespace {

constexpr uint64_t FS_FRONTIER_CHUNK = 1ull << 18;  // StreamVInt::MAX_INDEXES_COUNT
constexpr uint64_t FS_BUCKET_CHUNK = 1ull << 12;    // SmartMultiplexor small sector

using ps::GrowBuf;

// [0] in-segment children [1] count [2] written [3] cross children [4] children
// outside the segments [5] words scanned [6] frontier ids outside the segment
// [7] cross-segment slots filled [8] unsorted cross-segment bitmap chunks
constexpr int NCOUNTERS = 9;

uint64_t svi_worst(uint64_t m, bool bounds) {
    const uint64_t adj = (m + 7) / 8 * 8;
    return 4 + 17 * (adj / 4) + (bounds ? adj / 2 : 0);
}

}  // namespace

namespace psk {

// Column pass over the (workgroup, region) table: rcount[r] = marks of region
// r, and every entry replaced by the marks of the rows above it (the
// workgroup's first position in the region).  Workgroups of 16 waves take 64
// regions; wave w sums rows [32w, 32w + 32).
__global__ __launch_bounds__(FS_COL_WAVES * 64) void k_fs_bin_columns(uint32_t* table, uint32_t rows,
                                                                      uint32_t nregions, uint32_t* rcount) {
    __shared__ uint32_t part[FS_COL_WAVES][64];
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const uint32_t r = blockIdx.x * 64u + (uint32_t)lane;
    const bool ok = r < nregions;
    uint32_t v[FS_COL_ROWS];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < FS_COL_ROWS; ++k) {
        const uint32_t b = (uint32_t)(wave * FS_COL_ROWS + k);
        v[k] = (ok && b < rows) ? table[(uint64_t)b * nregions + r] : 0u;
        sum += v[k];
    }
    part[wave][lane] = sum;
    __syncthreads();
    uint32_t run = 0, total = 0;
    for (int w = 0; w < FS_COL_WAVES; ++w) {
        const uint32_t t = part[w][lane];
        run += (w < wave) ? t : 0u;
        total += t;
    }
#pragma unroll
    for (int k = 0; k < FS_COL_ROWS; ++k) {
        const uint32_t b = (uint32_t)(wave * FS_COL_ROWS + k);
        if (ok && b < rows) table[(uint64_t)b * nregions + r] = run;
        run += v[k];
    }
    if (ok && wave == 0) rcount[r] = total;
}

__global__ void k_fs_clear_odd(const uint32_t* idx, uint64_t n, uint64_t in_mask, unsigned long long* odd) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t id = idx[i];
        if ((uint64_t)id <= in_mask) odd[id >> 6] = 0ull;
    }
}

}  // namespace psk

struct ps_fs_segment {
    ps_puzzle puzzle{};
    ps_puzzle_info info{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int seg_bits = 0;
    uint32_t nseg = 0;
    int ops = 0, slots = 0, odd = 0;
    uint64_t nwords = 0, nregions = 0;
    int rb = 0;                                    // log2(states per region)
    uint32_t* rcount = nullptr;                    // marks per region
    unsigned long long* rstart = nullptr;          // nregions + 1
    uint32_t* cnt_all = nullptr;                   // entries per region
    unsigned long long* odd_bits = nullptr;
    unsigned long long* cnt_out = nullptr;
    unsigned long long* out_off = nullptr;
    unsigned long long* counters = nullptr;
    unsigned long long* h_counters = nullptr;  // pinned
    GrowBuf out_idx, out_ops, fr_bytes, slot_buf, compact, vals, xs_bytes, xseg_starts, block_cnt, block_off, marks,
        bin_table, fr_off, x_off;
    ps::GrowHost h_fr_off, h_x_off;  // chunk offsets of the frontier and cross-segment streams (pinned)
    uint32_t cus = 256;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // the cross-segment children are found, grouped and encoded on `side`
    // while the new frontier is encoded on `stream`
    hipStream_t side = nullptr;
    hipEvent_t ev_written = nullptr, ev_side = nullptr;
    // results of the last step (host)
    std::vector<uint64_t> fr_chunk_off;
    std::vector<uint32_t> bk_op, bk_seg;
    std::vector<uint64_t> bk_values, bk_byte_off;
    ps_fs_result last{};
    bool has_result = false;
};

namespace ps {
namespace {

int destroy_fs(ps_fs_segment* w) {
    if (!w) return PS_OK;
    if (w->stream) (void)hipStreamSynchronize(w->stream);
    for (GrowBuf* g : {&w->out_idx, &w->out_ops, &w->fr_bytes, &w->slot_buf, &w->compact, &w->vals, &w->xs_bytes,
                       &w->xseg_starts, &w->block_cnt, &w->block_off, &w->marks, &w->bin_table, &w->fr_off,
                       &w->x_off})
        g->release(w->stream);
    if (w->stream) (void)hipStreamSynchronize(w->stream);
    if (w->rcount) (void)hipFree(w->rcount);
    if (w->rstart) (void)hipFree(w->rstart);
    if (w->cnt_all) (void)hipFree(w->cnt_all);
    if (w->odd_bits) (void)hipFree(w->odd_bits);
    if (w->cnt_out) (void)hipFree(w->cnt_out);
    if (w->out_off) (void)hipFree(w->out_off);
    if (w->counters) (void)hipFree(w->counters);
    if (w->h_counters) (void)hipHostFree(w->h_counters);
    w->h_fr_off.release();
    w->h_x_off.release();
    if (w->side) (void)hipStreamSynchronize(w->side);
    if (w->ev0) (void)hipEventDestroy(w->ev0);
    if (w->ev1) (void)hipEventDestroy(w->ev1);
    if (w->ev_written) (void)hipEventDestroy(w->ev_written);
    if (w->ev_side) (void)hipEventDestroy(w->ev_side);
    if (w->side) (void)hipStreamDestroy(w->side);
    if (w->own_stream && w->stream) (void)hipStreamDestroy(w->stream);
    delete w;
    (void)hipGetLastError();
    return PS_OK;
}

int read_counters(ps_fs_segment* w, hipStream_t s) {
    PS_HIP_CHECK(hipMemcpyAsync(w->h_counters, w->counters, NCOUNTERS * 8, hipMemcpyDeviceToHost, s));
    return stream_wait(s);
}

// the region kernels need more than 64 KB of LDS for regions of 2^18 states
}  // namespace

int fs_bin_columns(uint32_t* table, uint32_t rows, uint32_t nregions, uint32_t* rcount, hipStream_t s) {
    if (rows > psk::FS_MAX_BIN_BLOCKS) return fail(PS_ERR_INTERNAL, "%u bin workgroups exceed %u", rows,
                                                   psk::FS_MAX_BIN_BLOCKS);
    hipLaunchKernelGGL(psk::k_fs_bin_columns, dim3((nregions + 63) / 64), dim3(psk::FS_COL_WAVES * 64), 0, s, table,
                       rows, nregions, rcount);
    PS_HIP_CHECK(hipGetLastError());
    return PS_OK;
}

namespace {

int region_lds_setup(size_t bytes) {
    static size_t set = 0;
    if (bytes <= 65536 || bytes <= set) return PS_OK;
    PS_HIP_CHECK(hipFuncSetAttribute((const void*)psk::k_fs_region<psk::FS_REGION_TPB, false>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    PS_HIP_CHECK(hipFuncSetAttribute((const void*)psk::k_fs_region<psk::FS_REGION_TPB, true>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    set = bytes;
    return PS_OK;
}

int run_step(ps_fs_segment* w, uint32_t segment, const uint32_t* d_idx, const uint8_t* d_ops, uint64_t n,
             const uint32_t* d_xidx, const uint64_t* xstarts, ps_fs_result* res) {
    hipStream_t s = w->stream;
    const auto t0 = std::chrono::steady_clock::now();
    w->has_result = false;
    PS_HIP_CHECK(hipEventRecord(w->ev0, s));
    if (int r = launch_fill(w->counters, 0, NCOUNTERS * 8, s)) return r;
    const uint64_t base = (uint64_t)segment << w->seg_bits;
    const uint64_t seg_mask = (1ull << w->seg_bits) - 1ull;
    const uint32_t grid_par = (uint32_t)std::min<uint64_t>(blocks_for(n, FS_TPB), (uint64_t)w->cus * 8);
    uint64_t bad_ids = 0;

    // 1 mark: in-segment children of the frontier and the incoming marks,
    //   counted per region, then placed in the mark list by region
    const uint64_t nx = xstarts ? xstarts[w->ops] : 0;
    const uint64_t max_marks = n * (uint64_t)w->slots + nx;
    if (max_marks > 0xFFFFFFFFull)
        return fail(PS_ERR_CAPACITY, "%llu marks in one step exceed 2^32", (unsigned long long)max_marks);
    psk::FsBinArgs a{};
    a.prm = params_of(w->puzzle);
    a.idx = d_idx;
    a.ops = d_ops;
    a.n = n;
    a.base = base;
    a.segment = segment;
    a.seg_bits = w->seg_bits;
    a.rb = w->rb;
    a.nregions = (uint32_t)w->nregions;
    a.nx = nx;
    a.nops = w->ops;
    if (nx) {
        if (int r = w->xseg_starts.ensure((size_t)(w->ops + 1) * 8, s)) return r;
        PS_HIP_CHECK(hipMemcpyAsync(w->xseg_starts.p, xstarts, (size_t)(w->ops + 1) * 8, hipMemcpyHostToDevice, s));
        a.xidx = d_xidx;
        a.xstarts = w->xseg_starts.as<uint64_t>();
    }
    a.odd = w->odd ? w->odd_bits : nullptr;
    a.children_out = w->counters + 0;
    a.bad_out = w->counters + 6;
    a.label_mask = 3u;
    // workgroups of the bin passes: two per CU, each with at least one pass of its threads
    const uint32_t grid_bin = (uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>({blocks_for(std::max(n, nx), psk::FS_BIN_TPB), (uint64_t)w->cus * 2,
                               (uint64_t)psk::FS_MAX_BIN_BLOCKS}));
    a.per = (n + grid_bin - 1) / grid_bin;
    a.perx = (nx + grid_bin - 1) / grid_bin;
    if (int r = w->bin_table.ensure((size_t)grid_bin * w->nregions * 4, s)) return r;
    a.table = w->bin_table.as<uint32_t>();
    if (int r = fs_bin(w->puzzle, a, 0, grid_bin, s)) return r;
    if (int r = fs_bin_columns(a.table, grid_bin, (uint32_t)w->nregions, w->rcount, s)) return r;
    if (int r = scan_excl_u32(w->rcount, w->nregions, w->rstart, s)) return r;
    if (int r = w->marks.ensure(std::max<uint64_t>(max_marks, 1) * 4, s)) return r;
    if (n || nx) {
        a.odd = nullptr;
        a.rstart = w->rstart;
        a.marks = w->marks.as<uint32_t>();
        if (int r = fs_bin(w->puzzle, a, 1, grid_bin, s)) return r;
    }

    // 2 scan: per region, the flags built in LDS from its marks and counted;
    //   an exclusive scan of the output counts; the ordered write
    psk::FsRegionArgs ra{};
    ra.nregions = (uint32_t)w->nregions;
    ra.marks = w->marks.as<uint32_t>();
    ra.rstart = w->rstart;
    ra.odd = w->odd ? w->odd_bits : nullptr;
    ra.rb = w->rb;
    ra.ops = w->ops;
    ra.cnt_out = w->cnt_out;
    ra.cnt_all = w->cnt_all;
    ra.out_off = w->out_off;
    const uint32_t rw = (1u << w->rb) / 8u;
    const size_t lds_region = ((size_t)rw + (rw + 31) / 32) * 4;  // flags and their nonzero-word bitmap
    if (int r = region_lds_setup(lds_region)) return r;
    const uint32_t grid_region = (uint32_t)std::min<uint64_t>(w->nregions, w->cus);
    hipLaunchKernelGGL((psk::k_fs_region<psk::FS_REGION_TPB, false>), dim3(grid_region),
                       dim3(psk::FS_REGION_TPB), lds_region, s, ra);
    PS_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(psk::k_fs_region_totals<1024>, dim3(1), dim3(1024), 0, s, (const uint32_t*)w->cnt_all,
                       (const unsigned long long*)w->rstart, (uint32_t)w->nregions, rw, w->counters);
    PS_HIP_CHECK(hipGetLastError());
    if (int r = scan_excl_u64(w->cnt_out, w->nregions, w->out_off, s)) return r;
    PS_HIP_CHECK(hipMemcpyAsync(w->counters + 2, w->out_off + w->nregions, 8, hipMemcpyDeviceToDevice, s));
    if (int r = read_counters(w, s)) return r;
    if (w->h_counters[6]) bad_ids = w->h_counters[6];
    const uint64_t children_in = w->h_counters[0];
    const uint64_t count = w->h_counters[1];
    const uint64_t n_out = w->h_counters[2];
    const uint64_t scanned = w->h_counters[5];
    if (int r = w->out_idx.ensure(std::max<uint64_t>(n_out, 1) * 4, s)) return r;
    if (int r = w->out_ops.ensure(std::max<uint64_t>(n_out, 1), s)) return r;
    ra.out_idx = w->out_idx.as<uint32_t>();
    ra.out_ops = w->out_ops.as<uint8_t>();
    hipLaunchKernelGGL((psk::k_fs_region<psk::FS_REGION_TPB, true>), dim3(grid_region),
                       dim3(psk::FS_REGION_TPB), lds_region, s, ra);
    PS_HIP_CHECK(hipGetLastError());
    if (w->odd && n) {
        hipLaunchKernelGGL(psk::k_fs_clear_odd, dim3(grid_par), dim3(256), 0, s, d_idx, n, seg_mask, w->odd_bits);
        PS_HIP_CHECK(hipGetLastError());
    }

    if (bad_ids) {
        PS_HIP_CHECK(hipStreamSynchronize(s));
        return fail(PS_ERR_INDEX_RANGE, "%llu frontier ids are not below 2^%d", (unsigned long long)bad_ids,
                    w->seg_bits);
    }

    // 3 cross-segment children of the written entries, found on the side
    //   stream while the main stream encodes the new frontier (4)
    const uint64_t nslots = n_out * (uint64_t)w->slots;
    if (nslots > 0x7FFFFFFFull)
        return fail(PS_ERR_CAPACITY, "%llu cross-segment slots in one step exceed 2^31",
                    (unsigned long long)nslots);
    hipStream_t s2 = w->side;
    PS_HIP_CHECK(hipEventRecord(w->ev_written, s));
    PS_HIP_CHECK(hipStreamWaitEvent(s2, w->ev_written, 0));
    uint64_t nblk = 0;
    if (n_out) {
        if (int r = w->slot_buf.ensure(nslots * 8, s2)) return r;
        psk::FsCrossArgs ca{};
        ca.prm = params_of(w->puzzle);
        ca.idx = w->out_idx.as<uint32_t>();
        ca.ops = w->out_ops.as<uint8_t>();
        ca.n = n_out;
        ca.base = base;
        ca.segment = segment;
        ca.nseg = w->nseg;
        ca.seg_bits = w->seg_bits;
        ca.slots = w->slot_buf.as<unsigned long long>();
        ca.counters = w->counters;
        nblk = (n_out + psk::FS_CROSS_PER_BLOCK - 1) / psk::FS_CROSS_PER_BLOCK;
        if (int r = w->block_cnt.ensure(nblk * 8, s2)) return r;
        if (int r = w->block_off.ensure((nblk + 1) * 8, s2)) return r;
        ca.block_count = w->block_cnt.as<unsigned long long>();
        if (int r = fs_cross(w->puzzle, ca, (uint32_t)nblk, s2, nullptr)) return r;
        if (int r = scan_excl_u64(w->block_cnt.as<unsigned long long>(), nblk, w->block_off.as<unsigned long long>(), s2))
            return r;
    }

    // 4 the new frontier stream (main stream); its chunk offsets come back with the step
    std::vector<uint64_t> starts;
    for (uint64_t c = 0; c < n_out; c += FS_FRONTIER_CHUNK) starts.push_back(c);
    starts.push_back(n_out);  // no chunk at all for an empty frontier
    uint64_t fr_cap = 0;
    for (size_t c = 0; c + 1 < starts.size(); ++c) fr_cap += svi_worst(starts[c + 1] - starts[c], true);
    const size_t nfo = starts.size();  // chunk offsets
    if (int r = w->fr_bytes.ensure(std::max<uint64_t>(fr_cap, 16), s)) return r;
    if (int r = w->fr_off.ensure(nfo * 8, s)) return r;
    if (int r = w->h_fr_off.ensure(nfo * 8)) return r;
    if (int r = svi_encode_chunks_dev(w->out_idx.as<uint32_t>(), w->out_ops.as<uint8_t>(), starts,
                                      w->fr_bytes.as<uint8_t>(), w->fr_bytes.cap, w->fr_off.as<uint64_t>(),
                                      nullptr, s))
        return r;
    PS_HIP_CHECK(hipMemcpyAsync(w->h_fr_off.p, w->fr_off.p, nfo * 8, hipMemcpyDeviceToHost, s));

    // 5 the cross-segment children compacted, bucketed by (op, segment) and
    //   encoded (side stream); the bucket byte offsets come back with the step
    uint64_t m = 0;
    if (n_out) {
        if (int r = read_counters(w, s2)) return r;
        if (w->h_counters[4])
            return fail(PS_ERR_INDEX_RANGE, "%llu children fall outside the %u segments",
                        (unsigned long long)w->h_counters[4], w->nseg);
        m = w->h_counters[3];
        if (w->h_counters[7] != m)
            return fail(PS_ERR_HIP, "%llu cross-segment slots filled for %llu cross-segment children",
                        (unsigned long long)w->h_counters[7], (unsigned long long)m);
        if (int r = w->compact.ensure(std::max<uint64_t>(m, 1) * 8, s2)) return r;
        hipLaunchKernelGGL(psk::k_fs_compact<256>, dim3((uint32_t)nblk), dim3(256), 0, s2,
                           (const unsigned long long*)w->slot_buf.as<unsigned long long>(),
                           (uint32_t)(psk::FS_CROSS_PER_BLOCK * w->slots),
                           (const unsigned long long*)w->block_cnt.as<unsigned long long>(),
                           (const unsigned long long*)w->block_off.as<unsigned long long>(),
                           w->compact.as<unsigned long long>());
        PS_HIP_CHECK(hipGetLastError());
    }
    w->bk_op.clear();
    w->bk_seg.clear();
    w->bk_values.clear();
    w->bk_byte_off.assign(1, 0);
    std::vector<uint64_t> first_chunk;
    uint64_t xbytes = 0, nchunks_x = 0, x_cap = 0;
    if (m) {
        // stable grouping by (op, segment): emission order inside a bucket
        if (int r = w->vals.ensure(m * 4 + 64, s2)) return r;
        std::vector<uint32_t> uk;
        std::vector<uint64_t> rc;
        uint64_t dropped = 0;
        if (int r = stable_group_by_key(w->compact.as<unsigned long long>(), m, (uint32_t)(w->ops * w->nseg),
                                        w->vals.as<uint32_t>(), uk, rc, &dropped, s2))
            return r;
        if (dropped) return fail(PS_ERR_INTERNAL, "%llu cross-segment keys out of range", (unsigned long long)dropped);
        const int nb = (int)uk.size();
        // 4096-value chunks inside every bucket
        std::vector<uint64_t> cstarts;
        first_chunk.assign(nb + 1, 0);
        uint64_t pos = 0;
        for (int b = 0; b < nb; ++b) {
            first_chunk[b] = cstarts.size();
            for (uint64_t c = 0; c < rc[b]; c += FS_BUCKET_CHUNK) {
                cstarts.push_back(pos + c);
                x_cap += svi_worst(std::min<uint64_t>(FS_BUCKET_CHUNK, rc[b] - c), false);
            }
            pos += rc[b];
        }
        first_chunk[nb] = cstarts.size();
        cstarts.push_back(pos);
        if (pos != m) return fail(PS_ERR_HIP, "bucket sizes %llu != %llu children", (unsigned long long)pos,
                                  (unsigned long long)m);
        nchunks_x = cstarts.size() - 1;
        if (int r = w->xs_bytes.ensure(x_cap + 16, s2)) return r;
        if (int r = w->x_off.ensure((nchunks_x + 1) * 8, s2)) return r;
        if (int r = w->h_x_off.ensure((nchunks_x + 1) * 8)) return r;
        if (int r = fc_encode_chunks_dev(w->vals.as<uint32_t>(), cstarts, 1, w->xs_bytes.as<uint8_t>(),
                                         w->xs_bytes.cap, w->x_off.as<uint64_t>(),
                                         reinterpret_cast<uint32_t*>(w->counters + 8), s2))
            return r;
        PS_HIP_CHECK(hipMemcpyAsync(w->h_x_off.p, w->x_off.p, (nchunks_x + 1) * 8, hipMemcpyDeviceToHost, s2));
        for (int b = 0; b < nb; ++b) {
            w->bk_op.push_back(uk[b] / w->nseg);
            w->bk_seg.push_back(uk[b] % w->nseg);
            w->bk_values.push_back(rc[b]);
        }
    }
    // the main stream waits for the side stream's work
    PS_HIP_CHECK(hipEventRecord(w->ev_side, s2));
    PS_HIP_CHECK(hipStreamWaitEvent(s, w->ev_side, 0));
    PS_HIP_CHECK(hipMemcpyAsync(w->h_counters + 8, w->counters + 8, 8, hipMemcpyDeviceToHost, s));
    PS_HIP_CHECK(hipEventRecord(w->ev1, s));
    if (int r = event_wait(w->ev1)) return r;
    float ms = 0.f;
    PS_HIP_CHECK(hipEventElapsedTime(&ms, w->ev0, w->ev1));
    const uint64_t* hfo = w->h_fr_off.as<uint64_t>();
    w->fr_chunk_off.assign(hfo, hfo + nfo);
    if (w->fr_chunk_off.back() > w->fr_bytes.cap)
        return fail(PS_ERR_INTERNAL, "frontier stream of %llu bytes overran its %zu-byte buffer",
                    (unsigned long long)w->fr_chunk_off.back(), w->fr_bytes.cap);
    if (nchunks_x) {
        const uint64_t* hxo = w->h_x_off.as<uint64_t>();
        xbytes = hxo[nchunks_x];
        if (xbytes > w->xs_bytes.cap)
            return fail(PS_ERR_INTERNAL, "cross-segment stream of %llu bytes overran its %zu-byte buffer",
                        (unsigned long long)xbytes, w->xs_bytes.cap);
        if (w->h_counters[8])
            return fail(PS_ERR_INTERNAL, "%llu cross-segment bitmap chunks are not sorted",
                        (unsigned long long)w->h_counters[8]);
        for (size_t b = 0; b + 1 < first_chunk.size(); ++b) w->bk_byte_off.push_back(hxo[first_chunk[b + 1]]);
    }

    ps_fs_result r{};
    r.count = count;
    r.frontier_count = n_out;
    r.frontier_bytes = w->fr_chunk_off.back();
    r.frontier_chunks = w->fr_chunk_off.size() - 1;
    r.children_in = children_in;
    r.children_cross = m;
    r.buckets = w->bk_op.size();
    r.cross_chunks = nchunks_x;
    r.cross_bytes = xbytes;
    r.scanned_words = scanned;
    // inputs + outputs + 8 B per child
    r.algorithmic_bytes = 5 * n + 4 * nx + 5 * n_out + r.frontier_bytes + xbytes + 8 * (children_in + nx + m);
    r.device_ms = ms;
    r.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    w->last = r;
    w->has_result = true;
    if (res) *res = r;
    return PS_OK;
}

}  // namespace
}  // namespace ps

using namespace ps;

extern "C" {

int ps_fs_segment_create(const ps_puzzle* p, int32_t segment_bits, ps_stream st, ps_fs_segment** out) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (!p || !out) return fail(PS_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    ps_puzzle_info info{};
    if (int r = puzzle_query(p, &info)) return r;
    if (info.operators_count < 1 || info.operators_count > 4)
        return fail(PS_ERR_UNSUPPORTED, "the segment step needs 1..4 moves (puzzle has %d)", info.operators_count);
    if (segment_bits < 6 || segment_bits > 32) return fail(PS_ERR_INVALID_ARGUMENT, "segment_bits must be 6..32");
    const uint64_t seg_size = 1ull << segment_bits;
    const uint64_t nseg64 = (info.indexes_count + seg_size - 1) / seg_size;
    if (nseg64 == 0 || nseg64 * (uint64_t)info.operators_count > (1ull << 26))  // (op, segment) bucket keys
        return fail(PS_ERR_INVALID_ARGUMENT, "%llu segments of 2^%d states are too many",
                    (unsigned long long)nseg64, segment_bits);
    ps_fs_segment* w = new (std::nothrow) ps_fs_segment();
    if (!w) return fail(PS_ERR_OUT_OF_MEMORY, "host allocation failed");
    auto cleanup = [&](int rc) {
        destroy_fs(w);
        return rc;
    };
    w->puzzle = *p;
    w->info = info;
    w->seg_bits = segment_bits;
    w->nseg = (uint32_t)nseg64;
    w->ops = info.operators_count;
    w->odd = info.has_odd_cycles;
    w->nwords = seg_size / 8;
    w->rb = std::min(psk::FS_BIN_BITS, segment_bits);
    w->nregions = seg_size >> w->rb;
    if (w->nregions > psk::FS_MAX_REGIONS) return cleanup(fail(PS_ERR_INTERNAL, "%llu regions exceed %u",
                                                              (unsigned long long)w->nregions, psk::FS_MAX_REGIONS));
    w->cus = (uint32_t)hip_cu_count();
    if (int r = fs_cross(*p, psk::FsCrossArgs{}, 0, nullptr, &w->slots)) return cleanup(r);
    if (st) {
        w->stream = stream_of(st);
    } else {
        if (hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking) != hipSuccess)
            return cleanup(fail(PS_ERR_HIP, "stream creation failed"));
        w->own_stream = true;
    }
    auto dalloc = [&](void** ptr, uint64_t bytes) -> int {
        hipError_t e = hipMalloc(ptr, bytes ? bytes : 16);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            *ptr = nullptr;
            return fail(PS_ERR_OUT_OF_MEMORY, "hipMalloc(%llu) failed: %s", (unsigned long long)bytes,
                        hipGetErrorString(e));
        }
        return PS_OK;
    };
    if (int r = dalloc((void**)&w->rcount, w->nregions * 4)) return cleanup(r);
    if (int r = dalloc((void**)&w->rstart, w->nregions * 8 + 8)) return cleanup(r);
    if (int r = dalloc((void**)&w->cnt_all, w->nregions * 4)) return cleanup(r);
    if (w->odd)
        if (int r = dalloc((void**)&w->odd_bits, (seg_size + 63) / 64 * 8)) return cleanup(r);
    if (int r = dalloc((void**)&w->cnt_out, w->nregions * 8)) return cleanup(r);
    if (int r = dalloc((void**)&w->out_off, w->nregions * 8 + 8)) return cleanup(r);
    if (int r = dalloc((void**)&w->counters, NCOUNTERS * 8)) return cleanup(r);
    if (hipHostMalloc((void**)&w->h_counters, NCOUNTERS * 8, hipHostMallocDefault) != hipSuccess)
        return cleanup(fail(PS_ERR_OUT_OF_MEMORY, "pinned host allocation failed"));
    if (hipEventCreate(&w->ev0) != hipSuccess || hipEventCreate(&w->ev1) != hipSuccess ||
        hipEventCreateWithFlags(&w->ev_written, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&w->ev_side, hipEventDisableTiming) != hipSuccess)
        return cleanup(fail(PS_ERR_HIP, "event creation failed"));
    if (hipStreamCreateWithFlags(&w->side, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(PS_ERR_HIP, "stream creation failed"));
    if (w->odd)
        if (int r = launch_fill(w->odd_bits, 0, (seg_size + 63) / 64 * 8, w->stream)) return cleanup(r);
    if (hipStreamSynchronize(w->stream) != hipSuccess) return cleanup(fail(PS_ERR_HIP, "stream synchronize failed"));
    *out = w;
    return PS_OK;
    PS_API_END
}

int ps_fs_segment_destroy(ps_fs_segment* w) {
    PS_API_BEGIN
    if (!w) return PS_OK;
    return destroy_fs(w);
    PS_API_END
}

int ps_fs_segment_info(const ps_fs_segment* w, uint32_t* segments, int32_t* operators, uint64_t* footprint_bytes) {
    PS_API_BEGIN
    if (!w) return fail(PS_ERR_INVALID_ARGUMENT, "NULL segment workspace");
    if (segments) *segments = w->nseg;
    if (operators) *operators = w->ops;
    if (footprint_bytes)
        *footprint_bytes = w->nregions * 32 + (w->odd ? ((1ull << w->seg_bits) + 63) / 64 * 8 : 0);
    return PS_OK;
    PS_API_END
}

int ps_fs_segment_step(ps_fs_segment* w, uint32_t segment, const uint32_t* d_idx, const uint8_t* d_ops, uint64_t n,
                       const uint32_t* d_xseg_idx, const uint64_t* xseg_op_starts, ps_fs_result* res) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (!w) return fail(PS_ERR_INVALID_ARGUMENT, "NULL segment workspace");
    if (segment >= w->nseg) return fail(PS_ERR_INDEX_RANGE, "segment %u outside [0, %u)", segment, w->nseg);
    if (n && (!d_idx || !d_ops)) return fail(PS_ERR_INVALID_ARGUMENT, "NULL frontier pointer");
    if (xseg_op_starts) {
        if (xseg_op_starts[0] != 0) return fail(PS_ERR_INVALID_ARGUMENT, "xseg_op_starts[0] must be 0");
        for (int op = 0; op < w->ops; ++op)
            if (xseg_op_starts[op + 1] < xseg_op_starts[op])
                return fail(PS_ERR_INVALID_ARGUMENT, "xseg_op_starts must not decrease");
        if (xseg_op_starts[w->ops] && !d_xseg_idx) return fail(PS_ERR_INVALID_ARGUMENT, "NULL cross-segment pointer");
    }
    return run_step(w, segment, d_idx, d_ops, n, d_xseg_idx, xseg_op_starts, res);
    PS_API_END
}

int ps_fs_segment_frontier(const ps_fs_segment* w, const uint32_t** d_idx, const uint8_t** d_ops,
                           const uint8_t** d_bytes, uint64_t* chunk_offsets, uint64_t cap) {
    PS_API_BEGIN
    if (!w) return fail(PS_ERR_INVALID_ARGUMENT, "NULL segment workspace");
    if (!w->has_result) return fail(PS_ERR_INVALID_ARGUMENT, "no completed step");
    if (d_idx) *d_idx = w->out_idx.ptr_or_null<uint32_t>();
    if (d_ops) *d_ops = w->out_ops.ptr_or_null<uint8_t>();
    if (d_bytes) *d_bytes = w->fr_bytes.ptr_or_null<uint8_t>();
    if (chunk_offsets) {
        if (cap < w->fr_chunk_off.size())
            return fail(PS_ERR_CAPACITY, "need %zu chunk offsets", w->fr_chunk_off.size());
        std::memcpy(chunk_offsets, w->fr_chunk_off.data(), w->fr_chunk_off.size() * 8);
    }
    return PS_OK;
    PS_API_END
}

int ps_fs_segment_buckets(const ps_fs_segment* w, uint32_t* op, uint32_t* segment, uint64_t* values,
                          uint64_t* byte_offsets, uint64_t cap, const uint8_t** d_bytes) {
    PS_API_BEGIN
    if (!w) return fail(PS_ERR_INVALID_ARGUMENT, "NULL segment workspace");
    if (!w->has_result) return fail(PS_ERR_INVALID_ARGUMENT, "no completed step");
    const size_t nb = w->bk_op.size();
    if (cap < nb) return fail(PS_ERR_CAPACITY, "need room for %zu buckets", nb);  // byte_offsets: cap + 1
    for (size_t b = 0; b < nb; ++b) {
        if (op) op[b] = w->bk_op[b];
        if (segment) segment[b] = w->bk_seg[b];
        if (values) values[b] = w->bk_values[b];
    }
    if (byte_offsets) std::memcpy(byte_offsets, w->bk_byte_off.data(), (nb + 1) * 8);
    if (d_bytes) *d_bytes = w->xs_bytes.ptr_or_null<uint8_t>();
    return PS_OK;
    PS_API_END
}

}  // extern "C"
