// The reference's device shim and expansion launchers (Puzzles/GPU.h,
// PancakeGpu.h, SlidingPuzzleGpu.h, HanoiTowersGPU.h) over the C ABI.  Every
// size argument is an element count, as in the reference; the byte sizes the
// ABI takes are computed here.
#include <cstdlib>
#include <stdexcept>
#include <string>

#include "GPU.h"
#include "HanoiTowersGPU.h"
#include "PancakeGpu.h"
#include "SlidingPuzzleGpu.h"
#include "Util.h"
#include "puzzlesearch_amd.h"

void ERR(int status) {
    if (status != PS_OK) {
        const std::string msg = ps_last_error();
        std::cerr << msg << std::endl;
        throw std::runtime_error(msg);
    }
}

namespace {

void Init() { EnsurePsRuntime(); }

template <class T>
T* CreateBuffer(int count) {
    Init();
    if (count < 0) throw std::runtime_error("CreateGPUBuffer: negative element count");
    void* p = nullptr;
    ERR(ps_alloc(&p, (uint64_t)count * sizeof(T)));
    return (T*)p;
}

void Expand(int kind, int size, int height, int flags, int mode, uint64_t* in, uint64_t* out, uint64_t count,
            CuStream stream) {
    Init();
    const ps_puzzle p{kind, size, height, flags};
    ERR(ps_expand(&p, mode, in, out, count, stream));
}

}  // namespace

// ------------------------------------------------------------ Puzzles/GPU.h

uint64_t* CreateGPUBuffer(int count) { return CreateBuffer<uint64_t>(count); }
uint32_t* CreateGPUBuffer32(int count) { return CreateBuffer<uint32_t>(count); }
void DestroyGPUBuffer(uint64_t* gpuBuffer) { ERR(ps_free(gpuBuffer)); }
void DestroyGPUBuffer32(uint32_t* gpuBuffer) { ERR(ps_free(gpuBuffer)); }

CuStream CreateCudaStream() {
    Init();
    ps_stream s = nullptr;
    ERR(ps_stream_create(&s));
    return s;
}

void DestroyCudaStream(CuStream stream) { ERR(ps_stream_destroy(stream)); }

void CopyToGpu(const uint64_t* buffer, uint64_t* gpuBuffer, size_t count, CuStream stream) {
    ERR(ps_memcpy_h2d_async(gpuBuffer, buffer, (uint64_t)count * sizeof(uint64_t), stream));
}

void CopyFromGpu(const uint64_t* gpuBuffer, uint64_t* buffer, size_t count, CuStream stream) {
    ERR(ps_memcpy_d2h_async(buffer, gpuBuffer, (uint64_t)count * sizeof(uint64_t), stream));
    ERR(ps_stream_sync(stream));
}

void CopyToGpu(const uint32_t* buffer, uint32_t* gpuBuffer, size_t count, CuStream stream) {
    ERR(ps_memcpy_h2d_async(gpuBuffer, buffer, (uint64_t)count * sizeof(uint32_t), stream));
}

void CopyFromGpu(const uint32_t* gpuBuffer, uint32_t* buffer, size_t count, CuStream stream) {
    ERR(ps_memcpy_d2h_async(buffer, gpuBuffer, (uint64_t)count * sizeof(uint32_t), stream));
    ERR(ps_stream_sync(stream));
}

// ------------------------------------------------------- Puzzles/PancakeGpu.h

void PancakeExpand(uint64_t* gpuIndexes, uint64_t* gpuExpanded, int size, uint64_t count, CuStream stream) {
    Expand(PS_PANCAKE, size, 0, 0, PS_EXPAND_FULL, gpuIndexes, gpuExpanded, count, stream);
}

void PancakeExpandInSegment(uint64_t* gpuIndexes, uint64_t* gpuExpanded, int size, uint64_t count, CuStream stream) {
    Expand(PS_PANCAKE, size, 0, 0, PS_EXPAND_IN_SEGMENT, gpuIndexes, gpuExpanded, count, stream);
}

void PancakeExpandCrossSegment(uint64_t* gpuIndexes, uint64_t* gpuExpanded, int size, bool invIndex,
                               uint64_t count, CuStream stream) {
    Expand(PS_PANCAKE, size, 0, invIndex ? PS_FLAG_INV_INDEX : 0, PS_EXPAND_CROSS_SEGMENT, gpuIndexes, gpuExpanded,
           count, stream);
}

void PancakeCrossSegmentPostProcessGPU(uint32_t* gpuIndexes, int segment, int size, uint64_t count,
                                       CuStream stream) {
    Init();
    ERR(ps_pancake_xseg_postprocess(size, segment, gpuIndexes, count, stream));
}

// ------------------------------------------------- Puzzles/SlidingPuzzleGpu.h

void GpuSlidingTilePuzzleSimpleExpand(uint64_t* gpuIndexes, uint64_t* gpuExpanded, int width, int size,
                                      uint64_t count, CuStream stream) {
    if (width <= 0 || size % width != 0)
        throw std::runtime_error("GpuSlidingTilePuzzleSimpleExpand: size is not a multiple of width");
    Expand(PS_SLIDING, width, size / width, 0, PS_EXPAND_FULL, gpuIndexes, gpuExpanded, count, stream);
}

template <int width, int height>
void GpuSlidingTilePuzzleOptimizedExpand(uint64_t* gpuIndexes, uint64_t* gpuExpanded, uint64_t count,
                                         CuStream stream) {
    Expand(PS_SLIDING_PACKED, width, height, 0, PS_EXPAND_FULL, gpuIndexes, gpuExpanded, count, stream);
}

template void GpuSlidingTilePuzzleOptimizedExpand<2, 2>(uint64_t*, uint64_t*, uint64_t, CuStream);
template void GpuSlidingTilePuzzleOptimizedExpand<3, 2>(uint64_t*, uint64_t*, uint64_t, CuStream);
template void GpuSlidingTilePuzzleOptimizedExpand<3, 3>(uint64_t*, uint64_t*, uint64_t, CuStream);
template void GpuSlidingTilePuzzleOptimizedExpand<4, 2>(uint64_t*, uint64_t*, uint64_t, CuStream);
template void GpuSlidingTilePuzzleOptimizedExpand<4, 3>(uint64_t*, uint64_t*, uint64_t, CuStream);
template void GpuSlidingTilePuzzleOptimizedExpand<4, 4>(uint64_t*, uint64_t*, uint64_t, CuStream);
template void GpuSlidingTilePuzzleOptimizedExpand<5, 2>(uint64_t*, uint64_t*, uint64_t, CuStream);
template void GpuSlidingTilePuzzleOptimizedExpand<5, 3>(uint64_t*, uint64_t*, uint64_t, CuStream);
template void GpuSlidingTilePuzzleOptimizedExpand<6, 2>(uint64_t*, uint64_t*, uint64_t, CuStream);
template void GpuSlidingTilePuzzleOptimizedExpand<7, 2>(uint64_t*, uint64_t*, uint64_t, CuStream);
template void GpuSlidingTilePuzzleOptimizedExpand<8, 2>(uint64_t*, uint64_t*, uint64_t, CuStream);

// --------------------------------------------------- Puzzles/HanoiTowersGPU.h

void GpuHanoiTowersExpand(uint64_t* gpuIndexes, uint64_t* gpuExpanded, int size, bool useSymmetry,
                          uint64_t count, CuStream stream) {
    Expand(PS_HANOI4, size, 0, useSymmetry ? PS_FLAG_SYMMETRY : 0, PS_EXPAND_FULL, gpuIndexes, gpuExpanded, count,
           stream);
}

void GpuHanoiTowersExpandOptimized(uint64_t* gpuIndexes, uint64_t* gpuExpanded, int size, bool useSymmetry,
                                   uint64_t count, CuStream stream) {
    Expand(PS_HANOI4_OPT, size, 0, useSymmetry ? PS_FLAG_SYMMETRY : 0, PS_EXPAND_FULL, gpuIndexes, gpuExpanded,
           count, stream);
}
