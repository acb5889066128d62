// 4-peg Hanoi expansion launchers with the reference's signatures
// (Puzzles/HanoiTowersGPU.h:5-19), each one ps_expand call on the caller's
// stream.  Six slots per parent: gpuExpanded[6i + k] = child << 4 | peg, or ~0.
#pragma once

#include "GPU.h"

// The reference's linked-list kernel layout (Puzzles/HanoiTowersGPU.cu:114-164).
void GpuHanoiTowersExpand(uint64_t* gpuIndexes, uint64_t* gpuExpanded, int size, bool useSymmetry,
                          uint64_t count, CuStream stream);

// The reference's bit-plane kernel layout (Puzzles/HanoiTowersGPU.cu:250-301).
void GpuHanoiTowersExpandOptimized(uint64_t* gpuIndexes, uint64_t* gpuExpanded, int size, bool useSymmetry,
                                   uint64_t count, CuStream stream);
