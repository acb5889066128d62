// Pancake expansion launchers with the reference's signatures
// (Puzzles/PancakeGpu.h:5-32), each one ps_expand / ps_pancake_xseg_postprocess
// call on the caller's stream.  Device I/O: gpuIndexes[i] = index << (size-1) |
// usedOpBits; gpuExpanded[i*S + k] = child << (size-1) | op, or ~0, with
// S = size-1 (PancakeExpand), 11 (InSegment) or size-12 (CrossSegment).
#pragma once

#include "GPU.h"

void PancakeExpand(uint64_t* gpuIndexes, uint64_t* gpuExpanded, int size, uint64_t count, CuStream stream);

// 29-bit segmented index, size > 12: flips of 2..12 elements (ops 0..10).
void PancakeExpandInSegment(uint64_t* gpuIndexes, uint64_t* gpuExpanded, int size, uint64_t count, CuStream stream);

// 29-bit segmented index, size > 12: flips of 13..size elements (ops 11..size-2),
// the child in the inverse-index form when invIndex.
void PancakeExpandCrossSegment(uint64_t* gpuIndexes, uint64_t* gpuExpanded, int size, bool invIndex,
                               uint64_t count, CuStream stream);

// Maps count in-segment ids of target `segment` back from the inverse-index
// form, in place (Puzzles/PancakeGpu.cu:179-193).
void PancakeCrossSegmentPostProcessGPU(uint32_t* gpuIndexes, int segment, int size, uint64_t count,
                                       CuStream stream);
