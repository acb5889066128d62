// Disk-based drivers with the reference's calling surface
// (Common/DiskBasedBFS.h:9-45): the segmented three-bit BFS over Stores,
// whose per-segment step runs on the device (ps_seg3_step).
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "Puzzle.h"
#include "Store.h"

// Common/DiskBasedBFS.h:9-14
struct PuzzleOptions {
    size_t maxSteps = 10000;
    int segmentBits = 32;
    // solver threads of the reference; the device runs the segments of a
    // layer one after the other on one stream, so this is accepted and unused
    int threads = 1;
    StoreOptions storeOptions;
};

// Common/DiskBasedBFS.h:16-36
struct SegmentedOptions {
    SegmentedOptions(class Puzzle& puzzle, PuzzleOptions& opts);
    Store MakeStore(std::string suffix);
    StoreSet MakeStoreSet(std::string suffix, int count);

    std::pair<int, uint32_t> GetSegIdx(uint64_t index) const {
        return {int(index >> Opts.segmentBits), uint32_t(index & SegmentMask)};
    }

    class Puzzle& Puzzle;
    PuzzleOptions Opts;
    uint64_t TotalSize;
    int Segments;
    uint64_t SegmentSize;
    uint64_t SegmentMask;
    int OperatorsCount;
    int OperatorsMask;
    bool HasOddLengthCycles;
};

// DiskBasedOptThreeBitBFS (Common/DiskBasedOptThreeBitBFS.cpp:164-252): per
// layer every segment's step reads its current and old frontier and the
// cross-segment marks bucketed for it from the Stores, runs on the device,
// and writes the new frontier and the buckets for the next layer; stops at an
// empty layer or after opts.maxSteps layers.  The puzzle must be a GPU plugin
// (PuzzleGpu).  Returns the layer counts.
std::vector<uint64_t> DiskBasedOptThreeBitBFS(Puzzle& puzzle, std::string initialState, PuzzleOptions opts = {});
