// Plugin interface of the search drivers, mirroring Common/Puzzle.h:11-120 of
// the reference (same class names, virtual functions, argument meanings and
// buffer sizes), so drivers written against the reference compile unchanged.
#pragma once

#include <atomic>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "Buffer.h"
#include "Util.h"

// Common/Puzzle.h:11-14
struct ExpandHint {
    int SegmentBits = 0;
    bool CrossSegment = false;
};

// Common/Puzzle.h:16-48
class Puzzle {
public:
    static constexpr size_t MAX_INDEXES_BUFFER = 2 * 1024 * 1024;
    static constexpr uint64_t INVALID_INDEX = uint64_t(-1);

public:
    virtual ~Puzzle() = default;

    virtual std::string Name() const = 0;
    virtual int OperatorsCount() const = 0;
    virtual int BranchingFactor() const = 0;
    virtual uint64_t IndexesCount() const = 0;
    virtual bool HasOddLengthCycles() const = 0;

    virtual std::string ToString(uint64_t index) = 0;
    virtual uint64_t Parse(std::string state) = 0;

    // Expands every parent; children of parent i go to the output vectors.
    // GPU puzzles keep INVALID_INDEX slots (the caller filters them).
    virtual void Expand(std::vector<uint64_t>& indexes, std::vector<int>& usedOperatorBits,
                        std::vector<uint64_t>& expandedIndexes, std::vector<int>& expandedOperators,
                        ExpandHint hint) = 0;

    // Common/Puzzle.h:39-44 (used by the segmented driver for cross-segment ids)
    virtual void CrossSegmentPostProcess(int op, int segment, int segmentBits, Buffer<uint32_t>& expandedIndexes) {
        (void)op;
        (void)segment;
        (void)segmentBits;
        (void)expandedIndexes;
    }

protected:
    // Common/Puzzle.cpp:14-26: reserve MAX_INDEXES_BUFFER * BranchingFactor(), clear.
    virtual void SetupOutputBuffers(std::vector<uint64_t>& expandedIndexes, std::vector<int>& expandedOperators) {
        const size_t cap = MAX_INDEXES_BUFFER * (size_t)BranchingFactor();
        if (expandedIndexes.capacity() < cap) expandedIndexes.reserve(cap);
        if (expandedOperators.capacity() < cap) expandedOperators.reserve(cap);
        expandedIndexes.clear();
        expandedOperators.clear();
    }
};

// Common/Puzzle.h:50-120: collects parents, expands them MAX_INDEXES_BUFFER at a
// time and calls func(child, op) for every child that is not INVALID_INDEX.
class ExpandBuffer {
public:
    explicit ExpandBuffer(Puzzle& puzzle) : puzzle(puzzle) {
        indexes.reserve(Puzzle::MAX_INDEXES_BUFFER);
        usedOperatorBits.reserve(Puzzle::MAX_INDEXES_BUFFER);
    }

    void SetExpandHint(int segmentBits, bool crossSegment) {
        hint.SegmentBits = segmentBits;
        hint.CrossSegment = crossSegment;
    }

    template <typename F>
    void Add(uint64_t index, int usedOperatorBits_, F func) {
        indexes.push_back(index);
        usedOperatorBits.push_back(usedOperatorBits_);
        if (indexes.size() == Puzzle::MAX_INDEXES_BUFFER) Expand(func);
    }

    template <typename F>
    void Finish(F func) {
        if (!indexes.empty()) Expand(func);
        SetExpandHint(0, false);
    }

    static uint64_t ExpandedNodes() { return s_nodes.load(); }
    static uint64_t ExpandedTimes() { return s_times.load(); }
    static uint64_t ExpandedNanos() { return s_nanos.load(); }
    static void PrintStats();

private:
    template <typename F>
    void Expand(F func) {
        Timer timer;
        puzzle.Expand(indexes, usedOperatorBits, childIndexes, childOperators, hint);
        s_times++;
        s_nodes += indexes.size();
        s_nanos += timer.Elapsed();
        for (size_t i = 0; i < childIndexes.size(); i++) {
            const uint64_t c = childIndexes[i];
            if (c != Puzzle::INVALID_INDEX) func(c, childOperators[i]);
        }
        indexes.clear();
        usedOperatorBits.clear();
        childIndexes.clear();
        childOperators.clear();
    }

    Puzzle& puzzle;
    ExpandHint hint;
    std::vector<uint64_t> indexes;
    std::vector<int> usedOperatorBits;
    std::vector<uint64_t> childIndexes;
    std::vector<int> childOperators;

    static std::atomic<uint64_t> s_nodes;
    static std::atomic<uint64_t> s_nanos;
    static std::atomic<uint64_t> s_times;
};
