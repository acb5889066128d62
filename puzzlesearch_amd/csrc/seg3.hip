// One disk-segment step of the three-bit BFS on the device:
// DB_Opt3BitBFS_Solver::Expand(segment) (Common/DiskBasedOptThreeBitBFS.cpp:54-144)
// with the segment's frontiers and incoming cross-segment marks already in HBM.
//
//   marks    the incoming cross-segment marks, after the 13..16-pancake
//            inverse-index post-process, set NextArray                 (:64-76)
//   expand   in-segment children of the current frontier set NextArray; with
//            odd cycles the frontier sets CurArray                     (:78-99)
//   clear    the old frontier clears NextArray                         (:101-111)
//   scan     ScanBitsAndClear[WithExcl CurArray] in two passes over the
//            touched regions: counts per region, a prefix sum, then the
//            ordered write of the new layer's ids and the clear    (:113-141)
//   stream   the new layer as FrontierCompression::Encode chunks of 2^18
//            values, as CompressedSegmentWriter writes it
//            (Common/CompressedFrontier.cpp:128-158)
//   cross    (a second stream, beside `stream`) cross-segment children of
//            the new layer in (parent, slot) order, stably grouped by
//            (op, segment), each bucket cut into 4096-value chunks compressed
//            with FrontierCompression::EncodeWithCheck, as SmartMultiplexor
//            flushes them (Common/Multiplexor.cpp:125-161)
//
// Both streams are encoded with the per-chunk choice, sizes and offsets on the
// device; the chunk offsets come back with the step's end.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <new>
#include <vector>

#include "codec_internal.h"
#include "devbuf.h"
#include "fs_kernels.h"
#include "fs_launch.h"
#include "ps_internal.h"
#include "seg3_launch.h"
#include "seg_sort.h"

namespace {

constexpr uint64_t S3_FRONTIER_CHUNK = 1ull << 18;  // StreamVInt::MAX_INDEXES_COUNT
constexpr uint64_t S3_BUCKET_CHUNK = 1ull << 12;    // SmartMultiplexor small sector

// FrontierCompression of m values is at most the StreamVInt size: a bitmap is
// chosen only when smaller
uint64_t fc_worst(uint64_t m) { return 4 + 17 * ((m + 7) / 8 * 2); }

}  // namespace

struct ps_seg3 {
    ps_puzzle puzzle{};
    ps_puzzle_info info{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int seg_bits = 0;
    uint32_t nseg = 0;
    int ops = 0, odd = 0, mode_in = PS_EXPAND_FULL, mode_x = PS_EXPAND_FULL, slots_x = 0;
    bool post = false;  // CrossSegmentPostProcess of the inverse-index pancake
    uint64_t nwords = 0, nregions = 0;
    unsigned long long* next = nullptr;
    unsigned long long* cur = nullptr;
    uint8_t* dirty = nullptr;
    unsigned long long* cnt_out = nullptr;
    unsigned long long* out_off = nullptr;
    unsigned long long* counters = nullptr;
    unsigned long long* h_counters = nullptr;  // pinned
    ps::GrowBuf out_idx, fr_bytes, xseg_copy, slot_buf, compact, vals, xs_bytes, block_cnt, block_off, fr_off, x_off;
    ps::GrowHost h_fr_off, h_x_off;  // chunk offsets of the frontier and cross-segment streams (pinned)
    // binned marks (puzzles expanded in full with at most 4 moves, no mark
    // post-process): per region of 2^brb states, no NextArray pass in HBM
    bool binned = false;
    int brb = 0;
    uint32_t bnreg = 0;
    uint32_t* brcount = nullptr;              // marks per region
    unsigned long long* brstart = nullptr;    // bnreg + 1
    unsigned long long* bostart = nullptr;    // bnreg + 1: ranges of the old frontier
    unsigned long long* bcnt_out = nullptr;   // bnreg
    unsigned long long* bout_off = nullptr;   // bnreg + 1
    ps::GrowBuf bin_table, bmarks;
    uint32_t cus = 256;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // the cross-segment children are found, grouped and encoded on `side`
    // while the new frontier is encoded on `stream`
    hipStream_t side = nullptr;
    hipEvent_t ev_written = nullptr, ev_side = nullptr;
    // results of the last step (host)
    std::vector<uint64_t> fr_chunk_off;
    std::vector<uint32_t> bk_op, bk_seg;
    std::vector<uint64_t> bk_values, bk_value_off, bk_byte_off;
    bool has_result = false;
};

namespace ps {
namespace {

int destroy_s3(ps_seg3* w) {
    if (!w) return PS_OK;
    if (w->stream) (void)hipStreamSynchronize(w->stream);
    if (w->side) (void)hipStreamSynchronize(w->side);
    for (GrowBuf* g : {&w->out_idx, &w->fr_bytes, &w->xseg_copy, &w->slot_buf, &w->compact, &w->vals, &w->xs_bytes,
                       &w->block_cnt, &w->block_off, &w->fr_off, &w->x_off, &w->bin_table, &w->bmarks})
        g->release(w->stream);
    if (w->stream) (void)hipStreamSynchronize(w->stream);
    w->h_fr_off.release();
    w->h_x_off.release();
    for (void* p : {(void*)w->next, (void*)w->cur, (void*)w->dirty, (void*)w->cnt_out, (void*)w->out_off,
                    (void*)w->counters, (void*)w->brcount, (void*)w->brstart, (void*)w->bostart, (void*)w->bcnt_out,
                    (void*)w->bout_off})
        if (p) (void)hipFree(p);
    if (w->h_counters) (void)hipHostFree(w->h_counters);
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

int read_counters(ps_seg3* w, hipStream_t s) {
    PS_HIP_CHECK(hipMemcpyAsync(w->h_counters, w->counters, psk::S3_NCOUNTERS * 8, hipMemcpyDeviceToHost, s));
    return stream_wait(s);
}

uint32_t grid_for(const ps_seg3* w, uint64_t n, uint32_t tpb) {
    return (uint32_t)std::min<uint64_t>(blocks_for(n, tpb), (uint64_t)w->cus * 8);
}

// Steps 1-4 with binned marks: the frontier's in-segment children and the
// incoming marks binned per region of 2^brb states (the frontier-search
// step's bin passes, labels dropped), then per region the NextArray bits in
// LDS, the old frontier cleared, the current frontier excluded (odd cycles;
// the count pass sets CurArray) and the new layer counted and written in
// order.  Returns with the counters read.
int binned_scan(ps_seg3* w, uint32_t segment, const uint32_t* d_cur, uint64_t ncur, const uint32_t* d_old,
                uint64_t nold, const uint32_t* d_xseg, const uint64_t* xstarts, uint64_t nx, hipStream_t s) {
    const uint64_t in_mask = (1ull << w->seg_bits) - 1ull;
    const uint64_t max_marks = ncur * (uint64_t)w->slots_x + nx;  // slots per parent in full expansion
    if (max_marks > 0xFFFFFFFFull)
        return fail(PS_ERR_CAPACITY, "%llu marks in one step exceed 2^32", (unsigned long long)max_marks);
    if (nx) {  // bad ids among the marks (the bin passes mask them)
        hipLaunchKernelGGL(psk::k_s3_check_list<256>, dim3(grid_for(w, nx, 256)), dim3(256), 0, s, d_xseg, nx, in_mask,
                           w->counters + 1, (unsigned long long*)nullptr);
        PS_HIP_CHECK(hipGetLastError());
    }
    if (nold) {  // the old frontier's ranges per region need it ascending
        hipLaunchKernelGGL(psk::k_s3_check_list<256>, dim3(grid_for(w, nold, 256)), dim3(256), 0, s, d_old, nold, in_mask,
                           (unsigned long long*)nullptr, w->counters + 8);
        PS_HIP_CHECK(hipGetLastError());
    }
    psk::FsBinArgs a{};
    a.prm = params_of(w->puzzle);
    a.idx = d_cur;
    a.ops = nullptr;
    a.n = ncur;
    a.base = (uint64_t)segment << w->seg_bits;
    a.segment = segment;
    a.seg_bits = w->seg_bits;
    a.rb = w->brb;
    a.nregions = w->bnreg;
    a.nx = nx;
    a.nops = w->ops;
    if (nx) {
        if (int r = w->xseg_copy.ensure((size_t)(w->ops + 1) * 8, s)) return r;
        if (int r = copy_to_device(w->xseg_copy.p, xstarts, (size_t)(w->ops + 1) * 8, s)) return r;
        a.xidx = d_xseg;
        a.xstarts = w->xseg_copy.as<uint64_t>();
    }
    a.odd = w->odd ? w->cur : nullptr;
    a.children_out = w->counters + 0;
    a.bad_out = w->counters + 1;
    a.label_mask = 0u;
    const uint32_t grid_bin = (uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>({blocks_for(std::max(ncur, nx), psk::FS_BIN_TPB), (uint64_t)w->cus * 2,
                               (uint64_t)psk::FS_MAX_BIN_BLOCKS}));
    a.per = (ncur + grid_bin - 1) / grid_bin;
    a.perx = (nx + grid_bin - 1) / grid_bin;
    if (int r = w->bin_table.ensure((size_t)grid_bin * w->bnreg * 4, s)) return r;
    a.table = w->bin_table.as<uint32_t>();
    if (int r = fs_bin(w->puzzle, a, 0, grid_bin, s)) return r;
    if (int r = fs_bin_columns(a.table, grid_bin, w->bnreg, w->brcount, s)) return r;
    if (int r = scan_excl_u32(w->brcount, w->bnreg, w->brstart, s)) return r;
    if (int r = w->bmarks.ensure(std::max<uint64_t>(max_marks, 1) * 4, s)) return r;
    a.odd = nullptr;
    a.rstart = w->brstart;
    a.marks = w->bmarks.as<uint32_t>();
    if (int r = fs_bin(w->puzzle, a, 1, grid_bin, s)) return r;
    // the old frontier's range in every region
    hipLaunchKernelGGL(psk::k_s3_range_starts<256>, dim3((w->bnreg + 1 + 255) / 256), dim3(256), 0, s, d_old, nold, w->brb,
                       w->bnreg, w->bostart);
    PS_HIP_CHECK(hipGetLastError());
    psk::S3RegionArgs ra{};
    ra.nregions = w->bnreg;
    ra.rb = w->brb;
    ra.marks = w->bmarks.as<uint32_t>();
    ra.rstart = w->brstart;
    ra.old = d_old;
    ra.ostart = w->bostart;
    ra.cur = w->odd ? w->cur : nullptr;
    ra.cnt_out = w->bcnt_out;
    ra.counters = w->counters;
    ra.out_off = w->bout_off;
    const uint32_t grid_region = (uint32_t)std::min<uint64_t>(w->bnreg, (uint64_t)w->cus * 2);
    hipLaunchKernelGGL((psk::k_s3_region<psk::S3_BIN_TPB, false>), dim3(grid_region), dim3(psk::S3_BIN_TPB), 0, s, ra);
    PS_HIP_CHECK(hipGetLastError());
    if (int r = scan_excl_u64(w->bcnt_out, w->bnreg, w->bout_off, s)) return r;
    if (int r = read_counters(w, s)) return r;
    if (w->h_counters[8]) {
        if (w->odd && ncur) {  // the count pass set CurArray: leave it clear
            hipLaunchKernelGGL(psk::k_s3_clear_cur<256>, dim3(grid_for(w, ncur, 256)), dim3(256), 0, s, d_cur, ncur,
                               in_mask, w->cur);
            PS_HIP_CHECK(hipGetLastError());
            PS_HIP_CHECK(hipStreamSynchronize(s));
        }
        return fail(PS_ERR_INVALID_ARGUMENT, "the old frontier is not ascending (%llu descents)",
                    (unsigned long long)w->h_counters[8]);
    }
    const uint64_t count = w->h_counters[2];
    if (int r = w->out_idx.ensure(std::max<uint64_t>(count, 1) * 4, s)) return r;
    ra.out_idx = w->out_idx.as<uint32_t>();
    hipLaunchKernelGGL((psk::k_s3_region<psk::S3_BIN_TPB, true>), dim3(grid_region), dim3(psk::S3_BIN_TPB),
                       (size_t)psk::S3_STAGE_IDS * 4, s, ra);
    PS_HIP_CHECK(hipGetLastError());
    return PS_OK;
}

int run_step(ps_seg3* w, uint32_t segment, const uint32_t* d_cur, uint64_t ncur, const uint32_t* d_old, uint64_t nold,
             const uint32_t* d_xseg, const uint64_t* xstarts, ps_seg3_result* res) {
    hipStream_t s = w->stream;
    const auto t0 = std::chrono::steady_clock::now();
    w->has_result = false;
    w->fr_chunk_off.assign(1, 0);
    w->bk_op.clear();
    w->bk_seg.clear();
    w->bk_values.clear();
    w->bk_value_off.assign(1, 0);
    w->bk_byte_off.assign(1, 0);
    const uint64_t nx = xstarts ? xstarts[w->ops] : 0;
    ps_seg3_result r{};
    if (nx == 0 && ncur == 0) {  // the segment has no data (:97)
        r.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        w->has_result = true;
        if (res) *res = r;
        return PS_OK;
    }
    PS_HIP_CHECK(hipEventRecord(w->ev0, s));
    if (int rc = launch_fill(w->counters, 0, psk::S3_NCOUNTERS * 8, s)) return rc;
    const uint64_t base = (uint64_t)segment << w->seg_bits;
    const uint64_t in_mask = (1ull << w->seg_bits) - 1ull;

    if (w->binned) {
        if (int rc = binned_scan(w, segment, d_cur, ncur, d_old, nold, d_xseg, xstarts, nx, s)) return rc;
    } else {
        // 1 incoming cross-segment marks
        if (nx) {
            const uint32_t* marks = d_xseg;
            if (w->post) {  // PancakeOptimizedGPU::CrossSegmentPostProcess on a copy
                if (int rc = w->xseg_copy.ensure(nx * 4, s)) return rc;
                PS_HIP_CHECK(hipMemcpyAsync(w->xseg_copy.p, d_xseg, nx * 4, hipMemcpyDeviceToDevice, s));
                if (int rc = ps_pancake_xseg_postprocess(w->puzzle.size, (int32_t)segment, w->xseg_copy.as<uint32_t>(), nx,
                                                         s))
                    return rc;
                marks = w->xseg_copy.as<uint32_t>();
            }
            hipLaunchKernelGGL(psk::k_s3_mark_list<256>, dim3(grid_for(w, nx, 256)), dim3(256), 0, s, marks, nx, in_mask,
                               (uint32_t*)w->next, w->dirty, w->counters + 1);
            PS_HIP_CHECK(hipGetLastError());
        }
        // 2 in-segment children of the current frontier
        if (ncur) {
            psk::S3MarkArgs a{};
            a.prm = params_of(w->puzzle);
            a.idx = d_cur;
            a.n = ncur;
            a.base = base;
            a.segment = segment;
            a.seg_bits = w->seg_bits;
            a.next = (uint32_t*)w->next;
            a.dirty = w->dirty;
            a.cur = w->odd ? w->cur : nullptr;
            a.counters = w->counters;
            if (int rc = s3_mark_in(w->puzzle, w->mode_in, a, grid_for(w, ncur, S3_TPB), s)) return rc;
        }
        // 3 the old frontier clears
        if (nold) {
            hipLaunchKernelGGL(psk::k_s3_clear_list<256>, dim3(grid_for(w, nold, 256)), dim3(256), 0, s, d_old, nold, in_mask,
                               (uint32_t*)w->next);
            PS_HIP_CHECK(hipGetLastError());
        }
        // 4 scan: count per region, prefix sum, ordered write + clear
        psk::S3ScanArgs sa{};
        sa.next = w->next;
        sa.cur = w->odd ? w->cur : nullptr;
        sa.dirty = w->dirty;
        sa.nwords = w->nwords;
        sa.cnt_out = w->cnt_out;
        sa.counters = w->counters;
        sa.out_off = w->out_off;
        const uint32_t grid_scan = (uint32_t)std::min<uint64_t>(w->nregions, (uint64_t)w->cus * 8);
        hipLaunchKernelGGL(psk::k_s3_count<S3_TPB>, dim3(grid_scan), dim3(S3_TPB), 0, s, sa);
        PS_HIP_CHECK(hipGetLastError());
        if (int rc = scan_excl_u64(w->cnt_out, w->nregions, w->out_off, s)) return rc;
        if (int rc = read_counters(w, s)) return rc;
        if (int rc = w->out_idx.ensure(std::max<uint64_t>(w->h_counters[2], 1) * 4, s)) return rc;
        sa.out_idx = w->out_idx.as<uint32_t>();
        hipLaunchKernelGGL(psk::k_s3_write<S3_TPB>, dim3(grid_scan), dim3(S3_TPB), 0, s, sa);
        PS_HIP_CHECK(hipGetLastError());
    }
    const uint64_t children_in = w->h_counters[0];
    const uint64_t bad_ids = w->h_counters[1];
    const uint64_t count = w->h_counters[2];
    const uint64_t scanned = w->h_counters[3];
    if (w->odd && ncur) {
        hipLaunchKernelGGL(psk::k_s3_clear_cur<256>, dim3(grid_for(w, ncur, 256)), dim3(256), 0, s, d_cur, ncur, in_mask,
                           w->cur);
        PS_HIP_CHECK(hipGetLastError());
    }
    if (bad_ids) {
        PS_HIP_CHECK(hipStreamSynchronize(s));
        return fail(PS_ERR_INDEX_RANGE, "%llu frontier or mark ids are not below 2^%d", (unsigned long long)bad_ids,
                    w->seg_bits);
    }

    // 5 cross-segment children of the new layer (side stream, beside 6)
    const uint64_t nslots = count * (uint64_t)w->slots_x;
    if (nslots > 0x7FFFFFFFull)
        return fail(PS_ERR_CAPACITY, "%llu cross-segment slots in one step exceed 2^31", (unsigned long long)nslots);
    hipStream_t s2 = w->side;
    PS_HIP_CHECK(hipEventRecord(w->ev_written, s));
    PS_HIP_CHECK(hipStreamWaitEvent(s2, w->ev_written, 0));
    uint64_t nblk = 0;
    if (count && w->slots_x) {
        if (int rc = w->slot_buf.ensure(nslots * 8, s2)) return rc;
        nblk = (count + psk::S3_CROSS_PER_BLOCK - 1) / psk::S3_CROSS_PER_BLOCK;
        if (int rc = w->block_cnt.ensure(nblk * 8, s2)) return rc;
        if (int rc = w->block_off.ensure((nblk + 1) * 8, s2)) return rc;
        psk::S3CrossArgs ca{};
        ca.prm = params_of(w->puzzle);
        ca.idx = w->out_idx.as<uint32_t>();
        ca.n = count;
        ca.base = base;
        ca.segment = segment;
        ca.nseg = w->nseg;
        ca.seg_bits = w->seg_bits;
        ca.slots = w->slot_buf.as<unsigned long long>();
        ca.counters = w->counters;
        ca.block_count = w->block_cnt.as<unsigned long long>();
        if (int rc = s3_cross(w->puzzle, w->mode_x, ca, (uint32_t)nblk, s2, nullptr)) return rc;
        if (int rc = scan_excl_u64(w->block_cnt.as<unsigned long long>(), nblk, w->block_off.as<unsigned long long>(),
                                   s2))
            return rc;
    }

    // 6 the new frontier stream (main stream); its chunk offsets come back with the step
    std::vector<uint64_t> starts;
    uint64_t fr_cap = 0;
    for (uint64_t c = 0; c < count; c += S3_FRONTIER_CHUNK) {
        starts.push_back(c);
        fr_cap += fc_worst(std::min<uint64_t>(S3_FRONTIER_CHUNK, count - c));
    }
    starts.push_back(count);  // no chunk at all for an empty layer
    const size_t nfo = starts.size();
    if (int rc = w->fr_bytes.ensure(std::max<uint64_t>(fr_cap, 16), s)) return rc;
    if (int rc = w->fr_off.ensure(nfo * 8, s)) return rc;
    if (int rc = w->h_fr_off.ensure(nfo * 8)) return rc;
    if (int rc = fc_encode_chunks_dev(w->out_idx.as<uint32_t>(), starts, 0, w->fr_bytes.as<uint8_t>(), w->fr_bytes.cap,
                                      w->fr_off.as<uint64_t>(), reinterpret_cast<uint32_t*>(w->counters + 7), s))
        return rc;
    PS_HIP_CHECK(hipMemcpyAsync(w->h_fr_off.p, w->fr_off.p, nfo * 8, hipMemcpyDeviceToHost, s));

    // 7 the cross-segment children compacted, grouped by (op, segment) and
    //   encoded (side stream); the bucket byte offsets come back with the step
    uint64_t m = 0, xbytes = 0, nchunks_x = 0;
    std::vector<uint64_t> first_chunk;
    const int cross_rc = [&]() -> int {
        if (count && w->slots_x) {
            if (int rc = read_counters(w, s2)) return rc;
            if (w->h_counters[5])
                return fail(PS_ERR_INDEX_RANGE, "%llu children fall outside the %u segments",
                            (unsigned long long)w->h_counters[5], w->nseg);
            m = w->h_counters[4];
            if (w->h_counters[6] != m)
                return fail(PS_ERR_INTERNAL, "%llu cross-segment children written for %llu counted",
                            (unsigned long long)w->h_counters[6], (unsigned long long)m);
            if (m) {
                if (int rc = w->compact.ensure(m * 8, s2)) return rc;
                hipLaunchKernelGGL(psk::k_fs_compact<256>, dim3((uint32_t)nblk), dim3(256), 0, s2,
                                   (const unsigned long long*)w->slot_buf.as<unsigned long long>(),
                                   (uint32_t)(psk::S3_CROSS_PER_BLOCK * w->slots_x),
                                   (const unsigned long long*)w->block_cnt.as<unsigned long long>(),
                                   (const unsigned long long*)w->block_off.as<unsigned long long>(),
                                   w->compact.as<unsigned long long>());
                PS_HIP_CHECK(hipGetLastError());
            }
        }
        if (!m) return PS_OK;
        if (int rc = w->vals.ensure(m * 4 + 64, s2)) return rc;
        std::vector<uint32_t> uk;
        std::vector<uint64_t> rc_;
        uint64_t dropped = 0;
        if (int rc = stable_group_by_key(w->compact.as<unsigned long long>(), m, (uint32_t)(w->ops * w->nseg),
                                         w->vals.as<uint32_t>(), uk, rc_, &dropped, s2))
            return rc;
        if (dropped) return fail(PS_ERR_INTERNAL, "%llu cross-segment keys out of range", (unsigned long long)dropped);
        std::vector<uint64_t> cstarts;
        first_chunk.assign(uk.size() + 1, 0);
        uint64_t pos = 0, cap = 0;
        for (size_t b = 0; b < uk.size(); ++b) {
            first_chunk[b] = cstarts.size();
            for (uint64_t c = 0; c < rc_[b]; c += S3_BUCKET_CHUNK) {
                cstarts.push_back(pos + c);
                cap += fc_worst(std::min<uint64_t>(S3_BUCKET_CHUNK, rc_[b] - c));
            }
            pos += rc_[b];
        }
        first_chunk[uk.size()] = cstarts.size();
        cstarts.push_back(pos);
        if (pos != m)
            return fail(PS_ERR_INTERNAL, "bucket sizes %llu != %llu children", (unsigned long long)pos,
                        (unsigned long long)m);
        nchunks_x = cstarts.size() - 1;
        if (int rc = w->xs_bytes.ensure(cap + 16, s2)) return rc;
        if (int rc = w->x_off.ensure((nchunks_x + 1) * 8, s2)) return rc;
        if (int rc = w->h_x_off.ensure((nchunks_x + 1) * 8)) return rc;
        if (int rc = fc_encode_chunks_dev(w->vals.as<uint32_t>(), cstarts, 1, w->xs_bytes.as<uint8_t>(),
                                          w->xs_bytes.cap, w->x_off.as<uint64_t>(),
                                          reinterpret_cast<uint32_t*>(w->counters + 7), s2))
            return rc;
        PS_HIP_CHECK(hipMemcpyAsync(w->h_x_off.p, w->x_off.p, (nchunks_x + 1) * 8, hipMemcpyDeviceToHost, s2));
        uint64_t vpos = 0;
        for (size_t b = 0; b < uk.size(); ++b) {
            w->bk_op.push_back(uk[b] / w->nseg);
            w->bk_seg.push_back(uk[b] % w->nseg);
            w->bk_values.push_back(rc_[b]);
            vpos += rc_[b];
            w->bk_value_off.push_back(vpos);
        }
        return PS_OK;
    }();
    if (cross_rc) {  // leave no work of this step queued behind the error
        (void)hipStreamSynchronize(s2);
        (void)hipStreamSynchronize(s);
        (void)hipGetLastError();
        return cross_rc;
    }
    // the main stream waits for the side stream's work
    PS_HIP_CHECK(hipEventRecord(w->ev_side, s2));
    PS_HIP_CHECK(hipStreamWaitEvent(s, w->ev_side, 0));
    PS_HIP_CHECK(hipMemcpyAsync(w->h_counters + 7, w->counters + 7, 8, hipMemcpyDeviceToHost, s));
    PS_HIP_CHECK(hipEventRecord(w->ev1, s));
    if (int r = event_wait(w->ev1)) return r;
    float ms = 0.f;
    PS_HIP_CHECK(hipEventElapsedTime(&ms, w->ev0, w->ev1));
    const uint64_t* hfo = w->h_fr_off.as<uint64_t>();
    w->fr_chunk_off.assign(hfo, hfo + nfo);
    if (w->fr_chunk_off.back() > w->fr_bytes.cap)
        return fail(PS_ERR_INTERNAL, "frontier stream of %llu bytes overran its %zu-byte buffer",
                    (unsigned long long)w->fr_chunk_off.back(), w->fr_bytes.cap);
    if (w->h_counters[7])
        return fail(PS_ERR_INVALID_ARGUMENT, "%llu frontier chunks are not sorted",
                    (unsigned long long)w->h_counters[7]);
    if (nchunks_x) {
        const uint64_t* hxo = w->h_x_off.as<uint64_t>();
        xbytes = hxo[nchunks_x];
        if (xbytes > w->xs_bytes.cap)
            return fail(PS_ERR_INTERNAL, "cross-segment stream of %llu bytes overran its %zu-byte buffer",
                        (unsigned long long)xbytes, w->xs_bytes.cap);
        for (size_t b = 0; b + 1 < first_chunk.size(); ++b) w->bk_byte_off.push_back(hxo[first_chunk[b + 1]]);
    }

    r.count = count;
    r.frontier_bytes = w->fr_chunk_off.back();
    r.frontier_chunks = w->fr_chunk_off.size() - 1;
    r.children_in = children_in;
    r.children_cross = m;
    r.buckets = w->bk_op.size();
    r.cross_chunks = nchunks_x;
    r.cross_bytes = xbytes;
    r.scanned_words = scanned;
    // inputs (ids and marks) + outputs (ids, streams, bucket entries) + 8 B per
    // child marked + the scan's read of next and cur and its clear of next
    r.algorithmic_bytes = 4 * (ncur + nold + nx) + 4 * count + r.frontier_bytes + 12 * m + xbytes +
                          8 * (children_in + nx) + 24 * scanned;
    r.device_ms = ms;
    r.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    w->has_result = true;
    if (res) *res = r;
    return PS_OK;
}

}  // namespace
}  // namespace ps

using namespace ps;

extern "C" {

int ps_seg3_create(const ps_puzzle* p, int32_t segment_bits, ps_stream st, ps_seg3** out) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (!p || !out) return fail(PS_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    ps_puzzle_info info{};
    if (int r = puzzle_query(p, &info)) return r;
    if (segment_bits < 1 || segment_bits > 32) return fail(PS_ERR_INVALID_ARGUMENT, "segment_bits must be 1..32");
    const uint64_t seg_size = 1ull << segment_bits;
    const uint64_t nseg64 = (info.indexes_count + seg_size - 1) / seg_size;
    if (nseg64 == 0 || nseg64 * (uint64_t)info.operators_count > (1ull << 26))  // (op, segment) bucket keys
        return fail(PS_ERR_INVALID_ARGUMENT, "%llu segments of 2^%d states are too many",
                    (unsigned long long)nseg64, segment_bits);
    ps_seg3* w = new (std::nothrow) ps_seg3();
    if (!w) return fail(PS_ERR_OUT_OF_MEMORY, "host allocation failed");
    auto cleanup = [&](int rc) {
        destroy_s3(w);
        return rc;
    };
    w->puzzle = *p;
    w->info = info;
    w->seg_bits = segment_bits;
    w->nseg = (uint32_t)nseg64;
    w->ops = info.operators_count;
    w->odd = info.has_odd_cycles;
    // the plugins' ExpandHint (Puzzles/PancakeOptimizedGpu.cpp:57-79)
    const bool split = p->kind == PS_PANCAKE && segment_bits == 29 && p->size > 12;
    w->mode_in = split ? PS_EXPAND_IN_SEGMENT : PS_EXPAND_FULL;
    w->mode_x = split ? PS_EXPAND_CROSS_SEGMENT : PS_EXPAND_FULL;
    w->post = split && (p->flags & PS_FLAG_INV_INDEX) != 0;
    w->nwords = (seg_size + 63) / 64;
    w->nregions = (seg_size + (1ull << psk::S3_REGION_BITS) - 1) >> psk::S3_REGION_BITS;
    w->cus = (uint32_t)hip_cu_count();
    // the puzzle has a device expander for this mode (and for mode_in, which
    // exists for every puzzle that has mode_x)
    if (int r = s3_cross(*p, w->mode_x, psk::S3CrossArgs{}, 0, nullptr, &w->slots_x)) return cleanup(r);
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
    // binned marks for puzzles expanded in full with at most 4 moves (the bin
    // passes of the frontier-search step) and no mark post-process
    w->binned = w->mode_in == PS_EXPAND_FULL && !w->post && w->ops <= 4 && segment_bits >= 6;
    if (w->binned) {
        w->brb = std::min(psk::FS_BIN_BITS, segment_bits);
        w->bnreg = (uint32_t)(seg_size >> w->brb);
        if (int r = dalloc((void**)&w->brcount, (uint64_t)w->bnreg * 4)) return cleanup(r);
        if (int r = dalloc((void**)&w->brstart, (uint64_t)w->bnreg * 8 + 8)) return cleanup(r);
        if (int r = dalloc((void**)&w->bostart, (uint64_t)w->bnreg * 8 + 8)) return cleanup(r);
        if (int r = dalloc((void**)&w->bcnt_out, (uint64_t)w->bnreg * 8)) return cleanup(r);
        if (int r = dalloc((void**)&w->bout_off, (uint64_t)w->bnreg * 8 + 8)) return cleanup(r);
    } else {
        if (int r = dalloc((void**)&w->next, w->nwords * 8)) return cleanup(r);
        if (int r = dalloc((void**)&w->dirty, w->nregions)) return cleanup(r);
        if (int r = dalloc((void**)&w->cnt_out, w->nregions * 8)) return cleanup(r);
        if (int r = dalloc((void**)&w->out_off, w->nregions * 8 + 8)) return cleanup(r);
    }
    if (w->odd)
        if (int r = dalloc((void**)&w->cur, w->nwords * 8)) return cleanup(r);
    if (int r = dalloc((void**)&w->counters, psk::S3_NCOUNTERS * 8)) return cleanup(r);
    if (hipHostMalloc((void**)&w->h_counters, psk::S3_NCOUNTERS * 8, hipHostMallocDefault) != hipSuccess)
        return cleanup(fail(PS_ERR_OUT_OF_MEMORY, "pinned host allocation failed"));
    if (hipEventCreate(&w->ev0) != hipSuccess || hipEventCreate(&w->ev1) != hipSuccess ||
        hipEventCreateWithFlags(&w->ev_written, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&w->ev_side, hipEventDisableTiming) != hipSuccess)
        return cleanup(fail(PS_ERR_HIP, "event creation failed"));
    if (hipStreamCreateWithFlags(&w->side, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(PS_ERR_HIP, "stream creation failed"));
    if (!w->binned) {
        if (int r = launch_fill(w->next, 0, w->nwords * 8, w->stream)) return cleanup(r);
        if (int r = launch_fill(w->dirty, 0, w->nregions, w->stream)) return cleanup(r);
    }
    if (w->odd)
        if (int r = launch_fill(w->cur, 0, w->nwords * 8, w->stream)) return cleanup(r);
    if (hipStreamSynchronize(w->stream) != hipSuccess) return cleanup(fail(PS_ERR_HIP, "stream synchronize failed"));
    *out = w;
    return PS_OK;
    PS_API_END
}

int ps_seg3_destroy(ps_seg3* w) {
    PS_API_BEGIN
    if (!w) return PS_OK;
    return destroy_s3(w);
    PS_API_END
}

int ps_seg3_info(const ps_seg3* w, uint32_t* segments, int32_t* operators, uint64_t* footprint_bytes) {
    PS_API_BEGIN
    if (!w) return fail(PS_ERR_INVALID_ARGUMENT, "NULL segment workspace");
    if (segments) *segments = w->nseg;
    if (operators) *operators = w->ops;
    if (footprint_bytes)
        *footprint_bytes = w->binned ? (uint64_t)w->bnreg * 36 + (w->odd ? w->nwords * 8 : 0)
                                     : w->nwords * 8 * (w->odd ? 2 : 1) + w->nregions * 17;
    return PS_OK;
    PS_API_END
}

int ps_seg3_step(ps_seg3* w, uint32_t segment, const uint32_t* d_cur, uint64_t ncur, const uint32_t* d_old,
                 uint64_t nold, const uint32_t* d_xseg, const uint64_t* xseg_op_starts, ps_seg3_result* res) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (!w) return fail(PS_ERR_INVALID_ARGUMENT, "NULL segment workspace");
    if (segment >= w->nseg) return fail(PS_ERR_INDEX_RANGE, "segment %u outside [0, %u)", segment, w->nseg);
    if ((ncur && !d_cur) || (nold && !d_old)) return fail(PS_ERR_INVALID_ARGUMENT, "NULL frontier pointer");
    if (xseg_op_starts) {
        if (xseg_op_starts[0] != 0) return fail(PS_ERR_INVALID_ARGUMENT, "xseg_op_starts[0] must be 0");
        for (int op = 0; op < w->ops; ++op)
            if (xseg_op_starts[op + 1] < xseg_op_starts[op])
                return fail(PS_ERR_INVALID_ARGUMENT, "xseg_op_starts must not decrease");
        if (xseg_op_starts[w->ops] && !d_xseg) return fail(PS_ERR_INVALID_ARGUMENT, "NULL cross-segment pointer");
    }
    return run_step(w, segment, d_cur, ncur, d_old, nold, d_xseg, xseg_op_starts, res);
    PS_API_END
}

int ps_seg3_frontier(const ps_seg3* w, const uint32_t** d_idx, const uint8_t** d_bytes, uint64_t* chunk_offsets,
                     uint64_t cap) {
    PS_API_BEGIN
    if (!w) return fail(PS_ERR_INVALID_ARGUMENT, "NULL segment workspace");
    if (!w->has_result) return fail(PS_ERR_INVALID_ARGUMENT, "no completed step");
    if (d_idx) *d_idx = w->out_idx.ptr_or_null<uint32_t>();
    if (d_bytes) *d_bytes = w->fr_bytes.ptr_or_null<uint8_t>();
    if (chunk_offsets) {
        if (cap < w->fr_chunk_off.size()) return fail(PS_ERR_CAPACITY, "need %zu chunk offsets", w->fr_chunk_off.size());
        std::memcpy(chunk_offsets, w->fr_chunk_off.data(), w->fr_chunk_off.size() * 8);
    }
    return PS_OK;
    PS_API_END
}

int ps_seg3_buckets(const ps_seg3* w, uint32_t* op, uint32_t* segment, uint64_t* values, uint64_t* value_offsets,
                    uint64_t* byte_offsets, uint64_t cap, const uint32_t** d_values, const uint8_t** d_bytes) {
    PS_API_BEGIN
    if (!w) return fail(PS_ERR_INVALID_ARGUMENT, "NULL segment workspace");
    if (!w->has_result) return fail(PS_ERR_INVALID_ARGUMENT, "no completed step");
    const size_t nb = w->bk_op.size();
    if (cap < nb) return fail(PS_ERR_CAPACITY, "need room for %zu buckets", nb);  // offsets: cap + 1
    for (size_t b = 0; b < nb; ++b) {
        if (op) op[b] = w->bk_op[b];
        if (segment) segment[b] = w->bk_seg[b];
        if (values) values[b] = w->bk_values[b];
    }
    if (value_offsets) std::memcpy(value_offsets, w->bk_value_off.data(), (nb + 1) * 8);
    if (byte_offsets) std::memcpy(byte_offsets, w->bk_byte_off.data(), (nb + 1) * 8);
    if (d_values) *d_values = w->vals.ptr_or_null<uint32_t>();
    if (d_bytes) *d_bytes = w->xs_bytes.ptr_or_null<uint8_t>();
    return PS_OK;
    PS_API_END
}

}  // extern "C"
