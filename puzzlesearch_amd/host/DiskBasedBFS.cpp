// DiskBasedOptThreeBitBFS over Stores with the segment step on the device.
#include "DiskBasedBFS.h"

#include <map>
#include <stdexcept>

#include "CompressedFrontier.h"
#include "PuzzleGpu.h"

using namespace fcstream;

// Common/DiskBasedBFS.cpp:3-17
SegmentedOptions::SegmentedOptions(class Puzzle& puzzle, PuzzleOptions& opts) : Puzzle(puzzle), Opts(opts) {
    TotalSize = puzzle.IndexesCount();
    SegmentSize = 1ull << opts.segmentBits;
    SegmentMask = SegmentSize - 1;
    Segments = int((TotalSize + SegmentSize - 1) / SegmentSize);
    if (Segments == 1 && SegmentSize > TotalSize) SegmentSize = TotalSize;  // SegmentMask stays valid
    OperatorsCount = puzzle.OperatorsCount();
    OperatorsMask = (1 << OperatorsCount) - 1;
    HasOddLengthCycles = puzzle.HasOddLengthCycles();
}

Store SegmentedOptions::MakeStore(std::string suffix) {
    return Store::CreateFileStore(Segments, suffix, Opts.storeOptions);
}

StoreSet SegmentedOptions::MakeStoreSet(std::string suffix, int count) {
    StoreSet set;
    for (int i = 0; i < count; i++) set.Stores.emplace_back(MakeStore(suffix + "-" + std::to_string(i)));
    return set;
}

namespace {

struct StreamHandle {
    ps_stream s = nullptr;
    StreamHandle() { PsCheck(ps_stream_create(&s), "ps_stream_create"); }
    ~StreamHandle() { ps_stream_destroy(s); }
};

struct Seg3Handle {
    ps_seg3* w = nullptr;
    ~Seg3Handle() { ps_seg3_destroy(w); }
};

// Every chunk the Store holds for the segment, appended to `bytes`; returns
// the number of values they hold.
uint64_t ReadSegmentStream(Store& store, int segment, std::vector<uint8_t>& bytes) {
    store.Rewind(segment);
    std::vector<uint8_t> piece(XSEG_PIECE);
    std::vector<uint64_t> off;
    uint64_t values = 0;
    for (;;) {
        const size_t n = store.Read(segment, piece.data(), piece.size());
        if (n == 0) break;
        values += ChunkBoundaries(piece.data(), n, off);
        bytes.insert(bytes.end(), piece.begin(), piece.begin() + (long)n);
    }
    return values;
}

// SmartMultiplexor's flush of one (op, segment) bucket held on the host:
// EncodeWithCheck chunks of 4096 values (Common/Multiplexor.cpp:135-161)
void WriteHostBucket(Store& store, int segment, const std::vector<uint32_t>& values, DeviceBytes& dev,
                     DeviceBytes& scratch, ps_stream s) {
    uint32_t* d = (uint32_t*)dev.Reserve(values.size() * 4 + 64);
    PsCheck(ps_memcpy_h2d_async(d, values.data(), values.size() * 4, s), "bucket upload");
    std::vector<uint8_t> bytes;
    std::vector<uint64_t> off;
    EncodeFromDevice(d, values.size(), 4096, true, bytes, off, scratch, s);
    WritePieces(store, segment, bytes.data(), off, XSEG_PIECE);
}

}  // namespace

std::vector<uint64_t> DiskBasedOptThreeBitBFS(Puzzle& puzzle, std::string initialState, PuzzleOptions opts) {
    auto* gpu = dynamic_cast<PuzzleGpu*>(&puzzle);
    if (!gpu) throw std::invalid_argument("DiskBasedOptThreeBitBFS runs its segment steps on the GPU: the puzzle "
                                          "must be a PuzzleGpu plugin");
    ensure(opts.segmentBits >= 1 && opts.segmentBits <= 32);
    EnsurePsRuntime();
    SegmentedOptions sopts(puzzle, opts);
    const int ops = sopts.OperatorsCount;

    Store oldStore = sopts.MakeStore("frontier1");
    Store curStore = sopts.MakeStore("frontier2");
    Store newStore = sopts.MakeStore("frontier3");
    StoreSet curXseg = sopts.MakeStoreSet("xseg1", ops);
    StoreSet nextXseg = sopts.MakeStoreSet("xseg2", ops);
    // :182-187
    auto swapStores = [&]() {
        oldStore.DeleteAll();
        std::swap(oldStore, curStore);
        std::swap(curStore, newStore);
        curXseg.Swap(nextXseg);
    };

    StreamHandle stream;
    Seg3Handle ws;
    const ps_puzzle desc = gpu->Descriptor();
    PsCheck(ps_seg3_create(&desc, opts.segmentBits, stream.s, &ws.w), "ps_seg3_create");
    DeviceBytes dCur, dOld, dX, dTmp, scratch;

    // SetInitialNode (:34-52): the start state, and its cross-segment children
    const uint64_t initial = puzzle.Parse(initialState);
    const auto [seg0, idx0] = sopts.GetSegIdx(initial);
    {
        CompressedSegmentWriter writer(newStore);
        writer.SetSegment(seg0);
        writer.Add(idx0);
        writer.Flush();
        std::map<std::pair<int, int>, std::vector<uint32_t>> buckets;
        ExpandBuffer expander(puzzle);
        expander.SetExpandHint(opts.segmentBits, true);
        auto cross = [&](uint64_t child, int op) {
            const auto [s, i] = sopts.GetSegIdx(child);
            if (s != seg0) buckets[{op, s}].push_back(i);
        };
        expander.Add(initial, 0, cross);
        expander.Finish(cross);
        for (const auto& [key, values] : buckets)
            WriteHostBucket(nextXseg.Stores[key.first], key.second, values, dTmp, scratch, stream.s);
    }
    swapStores();
    std::vector<uint64_t> result{1};

    std::vector<uint8_t> bytes, xbytes;
    std::vector<uint64_t> off;
    std::vector<uint32_t> bkOp, bkSeg;
    std::vector<uint64_t> bkValues, bkValueOff, bkByteOff;
    while (result.size() <= opts.maxSteps) {
        uint64_t total = 0;
        for (int s = 0; s < sopts.Segments; s++) {
            // the marks bucketed for this segment, op by op (:64-76)
            std::vector<uint64_t> xstarts(ops + 1, 0);
            xbytes.clear();
            for (int op = 0; op < ops; op++)
                xstarts[op + 1] = xstarts[op] + ReadSegmentStream(curXseg.Stores[op], s, xbytes);
            const uint64_t nx = DecodeToDevice(xbytes, dX, scratch, stream.s);
            ensure(nx == xstarts[ops]);
            bytes.clear();
            ReadSegmentStream(curStore, s, bytes);
            const uint64_t ncur = DecodeToDevice(bytes, dCur, scratch, stream.s);
            bytes.clear();
            ReadSegmentStream(oldStore, s, bytes);
            const uint64_t nold = DecodeToDevice(bytes, dOld, scratch, stream.s);
            ps_seg3_result res{};
            PsCheck(ps_seg3_step(ws.w, (uint32_t)s, (const uint32_t*)dCur.Ptr(), ncur, (const uint32_t*)dOld.Ptr(),
                                 nold, (const uint32_t*)dX.Ptr(), xstarts.data(), &res),
                    "ps_seg3_step");
            oldStore.Delete(s);  // :111
            curXseg.Delete(s);   // :77
            if (res.count == 0) continue;
            total += res.count;
            // the new frontier, as CompressedSegmentWriter stores it
            const uint8_t* dBytes = nullptr;
            off.assign(res.frontier_chunks + 1, 0);
            PsCheck(ps_seg3_frontier(ws.w, nullptr, &dBytes, off.data(), off.size()), "ps_seg3_frontier");
            bytes.resize(res.frontier_bytes);
            PsCheck(ps_memcpy_d2h_async(bytes.data(), dBytes, bytes.size(), stream.s), "frontier download");
            PsCheck(ps_stream_sync(stream.s), "stream synchronize");
            WritePieces(newStore, s, bytes.data(), off, FRONTIER_PIECE);
            // the buckets, as SmartMultiplexor stores them
            const size_t nb = res.buckets;
            if (nb == 0) continue;
            bkOp.resize(nb);
            bkSeg.resize(nb);
            bkValues.resize(nb);
            bkValueOff.resize(nb + 1);
            bkByteOff.resize(nb + 1);
            const uint32_t* dVals = nullptr;
            const uint8_t* dX8 = nullptr;
            PsCheck(ps_seg3_buckets(ws.w, bkOp.data(), bkSeg.data(), bkValues.data(), bkValueOff.data(),
                                    bkByteOff.data(), nb, &dVals, &dX8),
                    "ps_seg3_buckets");
            xbytes.resize(res.cross_bytes);
            PsCheck(ps_memcpy_d2h_async(xbytes.data(), dX8, xbytes.size(), stream.s), "bucket download");
            PsCheck(ps_stream_sync(stream.s), "stream synchronize");
            for (size_t b = 0; b < nb; b++) {
                const uint8_t* p = xbytes.data() + bkByteOff[b];
                ChunkBoundaries(p, bkByteOff[b + 1] - bkByteOff[b], off);
                WritePieces(nextXseg.Stores[bkOp[b]], (int)bkSeg[b], p, off, XSEG_PIECE);
            }
        }
        if (total == 0) break;
        result.push_back(total);
        swapStores();
    }
    oldStore.DeleteAll();
    curStore.DeleteAll();
    newStore.DeleteAll();
    curXseg.DeleteAll();
    nextXseg.DeleteAll();
    return result;
}
