// GPU plugin base class, mirroring Puzzles/PuzzleGpu.h:9-40: a pool of
// (stream, device input buffer, device output buffer) triples, the
// index<<ops|opBits packing, and a virtual ExpandGpu that subclasses implement
// with one launcher call (PancakeGpu.h, SlidingPuzzleGpu.h, HanoiTowersGPU.h).
// Buffers, streams and copies go through the element-count shim of GPU.h.
#pragma once

#include <memory>
#include <mutex>
#include <vector>

#include "GPU.h"
#include "Puzzle.h"
#include "puzzlesearch_amd.h"

class PuzzleGpu : public Puzzle {
protected:
    // Puzzles/PuzzleGpu.cpp:4-15: a stream, MAX_INDEXES_BUFFER parents in and
    // MAX_INDEXES_BUFFER * branchingFactor slots out.
    struct Exec {
        CuStream stream = nullptr;
        uint64_t* gpuSrc = nullptr;
        uint64_t* gpuDst = nullptr;
        explicit Exec(int branchingFactor);
        ~Exec();
        Exec(const Exec&) = delete;
        Exec& operator=(const Exec&) = delete;
    };

public:
    int BranchingFactor() const override { return OperatorsCount(); }

    void Expand(std::vector<uint64_t>& indexes, std::vector<int>& usedOperatorBits,
                std::vector<uint64_t>& expandedIndexes, std::vector<int>& expandedOperators,
                ExpandHint hint) override;

    // Descriptor of this puzzle for the C ABI (device BFS, segment step).
    virtual ps_puzzle Descriptor() const = 0;

protected:
    virtual void ExpandGpu(uint64_t* gpuIndexes, uint64_t* gpuExpanded, uint64_t count, CuStream stream) = 0;

    Exec* AquireStream();
    void ReleaseStream(Exec* stream);
    void CombineIndexAndOpBits(std::vector<uint64_t>& indexes, std::vector<int>& usedOperatorBits);
    void SplitIndexAndOps(std::vector<uint64_t>& expandedIndexes, std::vector<int>& expandedOperators);
    // Copy of the packed parents to an acquired Exec, `launch(src, dst, count,
    // stream)`, copy back of slots * count results, unpack.  The Exec is
    // released on every path, also when the launch throws.
    template <class Launch>
    void RunBatch(std::vector<uint64_t>& indexes, std::vector<uint64_t>& expandedIndexes,
                  std::vector<int>& expandedOperators, int slots, Launch&& launch);

private:
    std::mutex m_Mutex;
    std::vector<std::unique_ptr<Exec>> m_Streams;
    std::vector<Exec*> m_FreeStreams;
};

template <class Launch>
void PuzzleGpu::RunBatch(std::vector<uint64_t>& indexes, std::vector<uint64_t>& expandedIndexes,
                         std::vector<int>& expandedOperators, int slots, Launch&& launch) {
    ensure(indexes.size() <= MAX_INDEXES_BUFFER && slots <= BranchingFactor());
    expandedIndexes.resize(indexes.size() * (size_t)slots);
    expandedOperators.resize(indexes.size() * (size_t)slots);
    Exec* ex = AquireStream();
    try {
        CopyToGpu(indexes.data(), ex->gpuSrc, indexes.size(), ex->stream);
        launch(ex->gpuSrc, ex->gpuDst, (uint64_t)indexes.size(), ex->stream);
        CopyFromGpu(ex->gpuDst, expandedIndexes.data(), expandedIndexes.size(), ex->stream);
    } catch (...) {
        ReleaseStream(ex);
        throw;
    }
    ReleaseStream(ex);
    SplitIndexAndOps(expandedIndexes, expandedOperators);
}
