// GPU plugin base class, mirroring Puzzles/PuzzleGpu.h:9-40: a pool of
// (stream, device input buffer, device output buffer) triples, the
// index<<ops|opBits packing, and a virtual ExpandGpu that subclasses implement
// with one ps_expand call.  Device work goes through the C ABI only.
#pragma once

#include <memory>
#include <mutex>
#include <vector>

#include "Puzzle.h"
#include "puzzlesearch_amd.h"

using CuStream = void*;  // Puzzles/GPU.h:8 (a ps_stream / hipStream_t)

class PuzzleGpu : public Puzzle {
protected:
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

    // Descriptor of this puzzle for the C ABI (device BFS, expand).
    virtual ps_puzzle Descriptor() const = 0;

protected:
    virtual void ExpandGpu(uint64_t* gpuIndexes, uint64_t* gpuExpanded, uint64_t count, CuStream stream) = 0;

    Exec* AquireStream();
    void ReleaseStream(Exec* stream);
    void CombineIndexAndOpBits(std::vector<uint64_t>& indexes, std::vector<int>& usedOperatorBits);
    void SplitIndexAndOps(std::vector<uint64_t>& expandedIndexes, std::vector<int>& expandedOperators);
    // H2D of the packed parents, `launch`, D2H of slots*count results, unpack.
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
    expandedIndexes.resize(indexes.size() * (size_t)slots);
    expandedOperators.resize(indexes.size() * (size_t)slots);
    Exec* ex = AquireStream();
    try {
        PsCheck(ps_memcpy_h2d_async(ex->gpuSrc, indexes.data(), indexes.size() * 8, ex->stream), "copy to GPU");
        launch(ex->gpuSrc, ex->gpuDst, (uint64_t)indexes.size(), ex->stream);
        PsCheck(ps_memcpy_d2h_async(expandedIndexes.data(), ex->gpuDst, expandedIndexes.size() * 8, ex->stream),
                "copy from GPU");
        PsCheck(ps_stream_sync(ex->stream), "stream synchronize");
    } catch (...) {
        ReleaseStream(ex);
        throw;
    }
    ReleaseStream(ex);
    SplitIndexAndOps(expandedIndexes, expandedOperators);
}
