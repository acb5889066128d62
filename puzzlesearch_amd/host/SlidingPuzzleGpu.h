// Sliding-tile expansion launchers with the reference's signatures
// (Puzzles/SlidingPuzzleGpu.h:5-18), each one ps_expand call on the caller's
// stream.  Four slots per parent: gpuExpanded[4i + k] = child << 4 | op, or ~0.
#pragma once

#include "GPU.h"

// Lexicographic rank over size! permutations, size = width * height.
void GpuSlidingTilePuzzleSimpleExpand(uint64_t* gpuIndexes, uint64_t* gpuExpanded, int width, int size,
                                      uint64_t count, CuStream stream);

// The packed (segment << 32 | index) encoding; instantiated for the shapes of
// Puzzles/SlidingPuzzleGpu.cu:289-310: 2x2, 3x2, 3x3, 4x2, 4x3, 4x4, 5x2, 5x3,
// 6x2, 7x2, 8x2.
template <int width, int height>
void GpuSlidingTilePuzzleOptimizedExpand(uint64_t* gpuIndexes, uint64_t* gpuExpanded, uint64_t count,
                                         CuStream stream);
