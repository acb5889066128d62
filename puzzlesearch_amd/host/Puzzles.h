// GPU puzzle plugins with the reference's class names and constructor
// arguments (Puzzles/*Gpu*.h).  Parse/ToString run on the host; Expand runs one
// ps_expand launch per batch.
#pragma once

#include <string>
#include <vector>

#include "PuzzleGpu.h"

// Puzzles/PancakeOptimizedGpu.h:7-43
class PancakeOptimizedGPU : public PuzzleGpu {
public:
    PancakeOptimizedGPU(int size, bool invIndex);
    std::string Name() const override;
    int OperatorsCount() const override { return m_Size - 1; }
    int BranchingFactor() const override { return m_Size - 1; }
    uint64_t IndexesCount() const override;
    bool HasOddLengthCycles() const override { return true; }
    std::string ToString(uint64_t index) override;
    uint64_t Parse(std::string state) override;
    void Expand(std::vector<uint64_t>& indexes, std::vector<int>& usedOperatorBits,
                std::vector<uint64_t>& expandedIndexes, std::vector<int>& expandedOperators,
                ExpandHint hint) override;
    void CrossSegmentPostProcess(int op, int segment, int segmentBits, Buffer<uint32_t>& expandedIndexes) override;
    ps_puzzle Descriptor() const override;
    int Size() const { return m_Size; }
    bool InvIndex() const { return m_InvIndex; }

protected:
    void ExpandGpu(uint64_t* gpuIndexes, uint64_t* gpuExpanded, uint64_t count, CuStream stream) override;

private:
    int m_Size;
    bool m_InvIndex;
};

// Puzzles/SlidingTilePuzzleGpu.h:7-33 (lexicographic rank over (w*h)!)
class SlidingTilePuzzleGpu : public PuzzleGpu {
public:
    SlidingTilePuzzleGpu(int width, int height);
    std::string Name() const override;
    int OperatorsCount() const override { return 4; }
    uint64_t IndexesCount() const override;
    bool HasOddLengthCycles() const override { return false; }
    std::string ToString(uint64_t index) override;
    uint64_t Parse(std::string state) override;
    ps_puzzle Descriptor() const override;

protected:
    void ExpandGpu(uint64_t* gpuIndexes, uint64_t* gpuExpanded, uint64_t count, CuStream stream) override;

private:
    int m_Width, m_Height, m_Size;
};

// Puzzles/SlidingTilePuzzleOptimized.h:13-53 (packed 32-bit segment encoding)
template <int width, int height>
class SlidingTilePuzzleOptimized : public PuzzleGpu {
public:
    static constexpr int size = width * height;
    std::string Name() const override;
    int OperatorsCount() const override { return 4; }
    uint64_t IndexesCount() const override;
    bool HasOddLengthCycles() const override { return false; }
    std::string ToString(uint64_t index) override;
    uint64_t Parse(std::string state) override;
    void Expand(std::vector<uint64_t>& indexes, std::vector<int>& usedOperatorBits,
                std::vector<uint64_t>& expandedIndexes, std::vector<int>& expandedOperators,
                ExpandHint hint) override;
    ps_puzzle Descriptor() const override { return ps_puzzle{PS_SLIDING_PACKED, width, height, 0}; }

protected:
    void ExpandGpu(uint64_t* gpuIndexes, uint64_t* gpuExpanded, uint64_t count, CuStream stream) override;
};

// Puzzles/FourPegHanoiGPU.h and FourPegHanoiOptimizedGPU.h
class FourPegHanoiGPUBase : public PuzzleGpu {
public:
    FourPegHanoiGPUBase(int size, bool useSymmetry, bool optimized);
    std::string Name() const override;
    int OperatorsCount() const override { return 4; }
    int BranchingFactor() const override { return 6; }
    uint64_t IndexesCount() const override { return 1ull << (2 * m_Size); }
    bool HasOddLengthCycles() const override { return true; }
    std::string ToString(uint64_t index) override;
    uint64_t Parse(std::string state) override;
    void Expand(std::vector<uint64_t>& indexes, std::vector<int>& usedOperatorBits,
                std::vector<uint64_t>& expandedIndexes, std::vector<int>& expandedOperators,
                ExpandHint hint) override;
    ps_puzzle Descriptor() const override;
    int Size() const { return m_Size; }
    bool UseSymmetry() const { return m_UseSymmetry; }

protected:
    void ExpandGpu(uint64_t* gpuIndexes, uint64_t* gpuExpanded, uint64_t count, CuStream stream) override;

private:
    int m_Size;
    bool m_UseSymmetry;
    bool m_Optimized;
};

class FourPegHanoiGPU : public FourPegHanoiGPUBase {
public:
    FourPegHanoiGPU(int size, bool useSymmetry) : FourPegHanoiGPUBase(size, useSymmetry, false) {}
};

class FourPegHanoiOptimizedGPU : public FourPegHanoiGPUBase {
public:
    FourPegHanoiOptimizedGPU(int size, bool useSymmetry) : FourPegHanoiGPUBase(size, useSymmetry, true) {}
};

// Host permutation ranking (lexicographic; digit i = #smaller elements right of i).
namespace ps_host {
uint64_t PermutationRank(const int* arr, int size);
void PermutationUnrank(uint64_t index, int* arr, int size);
}  // namespace ps_host
