// HanoiTowers<size> and the GPU launchers of the HanoiTowers vertical solver
// with the reference's interface (HanoiTowers/HanoiTowers.h:9-70), over the C
// ABI.  A state index holds disk d's peg in bits 2d..2d+1; states are reduced
// by the symmetry of pegs 1..3 (ordered by bottom disk, largest first).
// Segments hold 2^32 states.  The expansions run on the device; ExpandInSegment
// has the GPU kernels' slot form, as the reference (USE_GPU) does.
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "Buffer.h"

template <int size>
class HanoiTowers {
public:
    static constexpr int Size = size;

    static uint64_t IndexesCount() { return 1ull << (size * 2); }

    static int MaxSegments() { return int((IndexesCount() + 0xFFFFFFFFull) >> 32); }

    // disks listed per peg from the top, 1-based, pegs separated by 0
    // (HanoiTowers/HanoiTowers.cpp:108-121, :139-150)
    static std::string ToString(uint64_t index);
    static uint64_t Parse(std::string stateStr);

    static std::pair<int, uint32_t> SplitIndex(uint64_t index) { return {int(index >> 32), uint32_t(index)}; }

    // six reduced children per index, appended (:152-172, :322-327)
    static void Expand(const std::vector<uint64_t>& indexes, std::vector<uint64_t>& children);

    // children = count * 6 (or count * 3 without the smallest disk's moves)
    // in-segment ids, the parent's own id where a child leaves the segment
    // (GpuExpandInSegment, HanoiTowers/HanoiExpandGPU.cu:231-275)
    static void ExpandInSegment(int segment, size_t count, const uint32_t* indexes, std::vector<uint32_t>& children);
    static void ExpandInSegmentWithoutSmallest(int segment, size_t count, const uint32_t* indexes,
                                               std::vector<uint32_t>& children);

    // the children outside the segment, appended in move order (:174-198, :353-358)
    static void ExpandCrossSegment(int segment, const std::vector<uint32_t>& indexes, std::vector<uint64_t>& children);

public:
    // at least three pegs hold one of the segment's 16 disks: no disk above
    // them can move (HanoiTowers/HanoiTowers.h:37-46)
    static bool InSegmentOnly(uint32_t index) {
        constexpr uint32_t MASK = 0x55555555u;
        const uint32_t p0 = index >> 1, p1 = index;
        const bool z0 = ((~p0 & ~p1) & MASK) != 0;
        const bool z1 = ((~p0 & p1) & MASK) != 0;
        const bool z2 = ((p0 & ~p1) & MASK) != 0;
        const bool z3 = ((p0 & p1) & MASK) != 0;
        return int(z0) + int(z1) + int(z2) + int(z3) >= 3;
    }

    // the unreduced states a reduced one stands for (:48-55)
    static int PegsMult(uint64_t index) {
        constexpr uint64_t BIT_MASK = 0x5555555555555555ull;
        const uint64_t p0 = index >> 1, p1 = index;
        const uint64_t pegs1 = (~p0 & p1) & BIT_MASK;
        const uint64_t pegs2 = (p0 & ~p1) & BIT_MASK;
        return 6 - (int(pegs2 == 0) * 3) - int(pegs1 == 0);
    }
};

// HanoiTowers/HanoiTowers.h:64-67, HanoiTowers/HanoiExpandGPU.cu:231-275
template <int size>
void GpuExpandInSegment(int segment, size_t count, const uint32_t* indexes, std::vector<uint32_t>& children);
template <int size>
void GpuExpandInSegmentWithoutSmallest(int segment, size_t count, const uint32_t* indexes,
                                       std::vector<uint32_t>& children);
