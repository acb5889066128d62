// Internal helpers shared by the HIP translation units of libpuzzlesearch_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "puzzlesearch_amd.h"

namespace ps {

// Records the message for ps_last_error() on this thread and returns code.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();
bool runtime_ready();
int require_runtime();
// compute units and LDS bytes per block of the current device
int hip_cu_count();
size_t hip_lds_limit();
// stream-ordered allocation from the library's scratch pool; scratch_free
// keeps small blocks for reuse on the same stream and hands the others to
// hipFreeAsync
int scratch_alloc(void** p, size_t bytes, hipStream_t s);
void scratch_free(void* p, size_t bytes, hipStream_t s);
// host-to-device copy in stream order; sources up to STAGE_MAX_BYTES go
// through pinned staging, so the call does not wait for a staging copy
constexpr size_t STAGE_MAX_BYTES = 4u << 20;
int copy_to_device(void* dst, const void* src, size_t bytes, hipStream_t s);

// Host-side puzzle facts (Common/Puzzle.h:22-26 for each reference class).
int puzzle_query(const ps_puzzle* p, ps_puzzle_info* out);
int slots_per_parent(const ps_puzzle* p, int mode);
// Default LDS window (states per tile) for the device BFS: children that stay
// within the parent's aligned window of this size are marked in LDS.
uint64_t puzzle_locality_block(const ps_puzzle* p);

inline hipStream_t stream_of(ps_stream s) { return reinterpret_cast<hipStream_t>(s); }

inline uint64_t factorial(int n) {
    uint64_t r = 1;
    for (int i = 2; i <= n; ++i) r *= (uint64_t)i;
    return r;
}

int launch_fill(void* dst, uint8_t value, uint64_t bytes, hipStream_t s);

// Looks up, once per device, the device address of the static scratch array
// of the allocation-free bit scan (bits.hip); called by ps_init with the
// device current.
int bits_scratch_init(int device);

}  // namespace ps

#define PS_HIP_CHECK(expr)                                                                               \
    do {                                                                                                 \
        hipError_t _e = (expr);                                                                          \
        if (_e != hipSuccess) {                                                                          \
            (void)hipGetLastError(); /* reported here; not again by a later call */                      \
            return ::ps::fail(PS_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, \
                              __LINE__);                                                                 \
        }                                                                                                \
    } while (0)

// Wraps a C-ABI body so no C++ exception crosses the boundary.
#define PS_API_BEGIN try {
#define PS_API_END                                                              \
    }                                                                           \
    catch (const std::bad_alloc&) {                                             \
        return ::ps::fail(PS_ERR_OUT_OF_MEMORY, "host allocation failed");      \
    }                                                                           \
    catch (const std::exception& e) {                                           \
        return ::ps::fail(PS_ERR_INVALID_ARGUMENT, "exception: %s", e.what());  \
    }                                                                           \
    catch (...) {                                                               \
        return ::ps::fail(PS_ERR_INVALID_ARGUMENT, "unknown exception");        \
    }
