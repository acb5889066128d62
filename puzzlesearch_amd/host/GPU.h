// Device runtime shim with the names, signatures and element-count semantics
// of the reference's Puzzles/GPU.h:8-27, implemented over the C ABI
// (include/puzzlesearch_amd.h).  A plugin source written against the
// reference's shim (PuzzleGpu.cpp, PancakeOptimizedGpu.cpp, ...) compiles and
// runs against this header unchanged.
//
// Differences, all at the error path: ERR takes a ps_* status code instead of
// a cudaError_t, and throws std::runtime_error carrying ps_last_error(), as the
// reference's ERR throws with cudaGetErrorString (Puzzles/GPU.cu:5-10).  The
// library is initialised on the first call (device $PS_DEVICE, default 0) if
// the caller has not called ps_init, as the CUDA runtime initialises lazily.
// The reference's __device__ rank/unrank helpers (Puzzles/GPU.h:26-32) have no
// host counterpart here: the device kernels behind the launchers implement
// their own ranking.
#pragma once

#include <cstddef>
#include <cstdint>

using CuStream = void*;  // Puzzles/GPU.h:8: a hipStream_t (ps_stream)

// Throws std::runtime_error if status != PS_OK.
void ERR(int status);

// Device buffers of `count` elements (Puzzles/GPU.cu:12-30).
uint64_t* CreateGPUBuffer(int count);
uint32_t* CreateGPUBuffer32(int count);
void DestroyGPUBuffer(uint64_t* gpuBuffer);
void DestroyGPUBuffer32(uint32_t* gpuBuffer);

// Streams (Puzzles/GPU.cu:32-40).
CuStream CreateCudaStream();
void DestroyCudaStream(CuStream stream);

// `count` elements host -> device (asynchronous, stream order) and device ->
// host (waits for the stream, so the host buffer holds the data on return),
// Puzzles/GPU.cu:42-60.
void CopyToGpu(const uint64_t* buffer, uint64_t* gpuBuffer, size_t count, CuStream stream);
void CopyFromGpu(const uint64_t* gpuBuffer, uint64_t* buffer, size_t count, CuStream stream);
void CopyToGpu(const uint32_t* buffer, uint32_t* gpuBuffer, size_t count, CuStream stream);
void CopyFromGpu(const uint32_t* gpuBuffer, uint32_t* buffer, size_t count, CuStream stream);
