// Search drivers with the reference signatures (Common/InMemoryBFS.h:5-11).
//
// InMemoryTwoBitBFS / InMemoryThreeBitBFS: the reference algorithms on host
// bit arrays, expanding through Puzzle::Expand via ExpandBuffer -- with a
// PuzzleGpu plugin every batch is expanded on the GPU (the drop-in path).
//
// GpuInMemoryTwoBitBFS / GpuInMemoryThreeBitBFS: the same searches with the
// bit arrays and the whole layer loop resident on the GPU (ps_bfs_*).
#pragma once

#include <string>
#include <vector>

#include "PuzzleGpu.h"

std::vector<uint64_t> InMemoryTwoBitBFS(Puzzle& puzzle, std::string initialState);
std::vector<uint64_t> InMemoryThreeBitBFS(Puzzle& puzzle, std::string initialState);

struct GpuSearchOptions {
    int segmentBits = 0;       // 0: one segment
    bool batched = true;       // persistent kernels + hipGraph layer loop; false: per-segment launches
    size_t maxSteps = 0;       // 0: run to exhaustion
    uint64_t indexCount = 0;   // 0: IndexesCount()
    std::vector<uint64_t>* frontierHashes = nullptr;  // per-layer frontier hashes (optional)
    ps_bfs_stats* stats = nullptr;                    // run statistics (optional)
};

std::vector<uint64_t> GpuInMemoryThreeBitBFS(PuzzleGpu& puzzle, std::string initialState,
                                             const GpuSearchOptions& opts = {});
std::vector<uint64_t> GpuInMemoryTwoBitBFS(PuzzleGpu& puzzle, std::string initialState,
                                           const GpuSearchOptions& opts = {});
