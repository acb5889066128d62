// Plain batch expansion: the device side of PuzzleGpu::Expand
// (Puzzles/PuzzleGpu.cpp:53-82).  One thread per parent; the parent's slots
// are staged in LDS and written out with coalesced stores.
#include <cstring>
#include <stdexcept>

#include "dispatch.h"
#include "ps_internal.h"

namespace ps {

namespace {

bool packed_shape_supported(int w, int h) {
    static const int shapes[][2] = {{2, 2}, {3, 2}, {4, 2}, {5, 2}, {6, 2}, {7, 2}, {8, 2}, {3, 3}, {4, 3}, {5, 3}, {4, 4}};
    for (auto& s : shapes)
        if (s[0] == w && s[1] == h) return true;
    return false;
}

}  // namespace

int puzzle_query(const ps_puzzle* p, ps_puzzle_info* out) {
    if (!p || !out) return fail(PS_ERR_INVALID_ARGUMENT, "NULL argument");
    std::memset(out, 0, sizeof(*out));
    switch (p->kind) {
    case PS_PANCAKE: {
        // Puzzles/PancakeOptimized.h:14-20, PancakeOptimized.cpp:94-105
        if (p->size < 2 || p->size > 16) return fail(PS_ERR_INVALID_ARGUMENT, "pancake size %d not in [2,16]", p->size);
        out->operators_count = p->size - 1;
        out->branching_factor = p->size - 1;
        out->has_odd_cycles = 1;
        if (p->size < 12) {
            out->indexes_count = factorial(p->size);
        } else {
            uint64_t r = 1;
            for (int i = 13; i <= p->size; ++i) r *= (uint64_t)i;
            out->indexes_count = r << 29;
        }
        return PS_OK;
    }
    case PS_SLIDING: {
        // Puzzles/SlidingTilePuzzleGpu.h:13-24, SlidingTilePuzzleGpu.cpp:23-27
        const int n = p->size * p->height;
        if (p->size < 1 || p->height < 1 || n < 2 || n > 16)
            return fail(PS_ERR_INVALID_ARGUMENT, "sliding %dx%d not supported (2..16 cells)", p->size, p->height);
        out->operators_count = 4;
        out->branching_factor = 4;
        out->has_odd_cycles = 0;
        out->indexes_count = factorial(n);
        return PS_OK;
    }
    case PS_SLIDING_PACKED: {
        // Puzzles/SlidingTilePuzzleOptimized.cpp:12-26, 271-277
        if (!packed_shape_supported(p->size, p->height))
            return fail(PS_ERR_INVALID_ARGUMENT, "packed sliding %dx%d is not an instantiated shape", p->size, p->height);
        const int n = p->size * p->height;
        out->operators_count = 4;
        out->branching_factor = 4;
        out->has_odd_cycles = 0;
        const uint64_t segs = n == 16 ? 0xEDC + 1 : n == 15 ? 0xDC + 1 : n == 14 ? 0xC + 1 : 1;
        if (segs > 1) {
            out->indexes_count = segs << 32;
        } else {
            uint64_t m = 16;
            for (int i = 3; i < n; ++i) m *= (uint64_t)i;
            out->indexes_count = m;
        }
        return PS_OK;
    }
    case PS_HANOI4:
    case PS_HANOI4_OPT:
        // Puzzles/FourPegHanoiGPU.h:13-24 (size <= 24)
        if (p->size < 1 || p->size > 24) return fail(PS_ERR_INVALID_ARGUMENT, "hanoi disks %d not in [1,24]", p->size);
        out->operators_count = 4;
        out->branching_factor = 6;
        out->has_odd_cycles = 1;
        out->indexes_count = 1ull << (2 * p->size);
        return PS_OK;
    }
    return fail(PS_ERR_INVALID_ARGUMENT, "unknown puzzle kind %d", p->kind);
}

int slots_per_parent(const ps_puzzle* p, int mode) {
    ps_puzzle_info info;
    if (puzzle_query(p, &info) != PS_OK) return -1;
    if (p->kind == PS_PANCAKE) {
        if (mode == PS_EXPAND_FULL) return p->size - 1;
        if (p->size <= 12) return -1;
        if (mode == PS_EXPAND_IN_SEGMENT) return 11;
        if (mode == PS_EXPAND_CROSS_SEGMENT) return p->size - 12;
        return -1;
    }
    return mode == PS_EXPAND_FULL ? info.branching_factor : -1;
}

uint64_t puzzle_locality_block(const ps_puzzle* p) {
    ps_puzzle_info info;
    if (puzzle_query(p, &info) != PS_OK) return 64;
    const uint64_t n = info.indexes_count;
    switch (p->kind) {
    case PS_PANCAKE:
    case PS_SLIDING: {
        // flips of <= k elements (moves at positions >= N-k) stay in the
        // parent's k!-aligned block; 9! bits = 45360 bytes of LDS
        const int cells = p->kind == PS_PANCAKE ? p->size : p->size * p->height;
        if (cells >= 11) return factorial(9);
        if (cells == 10) return factorial(8);
        return 4096;
    }
    case PS_HANOI4:
    case PS_HANOI4_OPT: {
        // moves of disks < k change only the low 2k index bits
        uint64_t b = 1ull << 18;
        while (b > 4096 && n / b < 256) b >>= 2;
        return b;
    }
    case PS_SLIDING_PACKED:
        return 1ull << 16;
    }
    return 4096;
}

template <class PZ, int TPB>
__global__ __launch_bounds__(TPB) void k_expand_slots(const uint64_t* __restrict__ in, uint64_t* __restrict__ out,
                                                      uint64_t count, int ops, psd::PuzzleParams prm) {
    constexpr int S = PZ::SLOTS;
    __shared__ uint64_t stage[TPB * S];
    const uint64_t first = (uint64_t)blockIdx.x * TPB;
    const uint64_t i = first + threadIdx.x;
    uint64_t res[S];
#pragma unroll
    for (int k = 0; k < S; ++k) res[k] = psd::kInvalid;
    if (i < count) {
        const uint64_t v = in[i];
        const uint32_t op_bits = (uint32_t)(v & ((1ull << ops) - 1ull));
        const uint64_t index = v >> ops;
        PZ::expand(index, op_bits, prm,
                   [&](int slot, uint64_t child, uint32_t label) { res[slot] = (child << ops) | label; });
    }
#pragma unroll
    for (int k = 0; k < S; ++k) stage[threadIdx.x * S + k] = res[k];
    __syncthreads();
    const uint64_t n = (count - first < (uint64_t)TPB ? count - first : (uint64_t)TPB) * S;
    uint64_t* dst = out + first * S;
    for (uint64_t j = threadIdx.x; j < n; j += TPB) dst[j] = stage[j];
}

// PancakeCrossSegmentPostProcessGPU (Puzzles/PancakeGpu.cu:179-193)
template <int N>
__global__ void k_pancake_xseg_post(uint32_t* idx, uint32_t segment, uint64_t count) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint64_t index = ((uint64_t)segment << 29) | idx[i];
    uint32_t d[N], a[N];
    psd::pancake_digits<N>(index, d);
    psd::digits_to_perm<N>(d, a);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const uint32_t t = a[N - 12 + k];
        a[N - 12 + k] = a[N - 1 - k];
        a[N - 1 - k] = t;
    }
    psd::perm_to_digits<N>(a, d);
    uint64_t r = psd::pancake_index_from_digits<N>(d);
    if ((r >> 29) != segment) r = psd::kInvalid;
    idx[i] = (uint32_t)(r & ((1ull << 29) - 1ull));
}

}  // namespace ps

using namespace ps;

extern "C" int ps_expand(const ps_puzzle* p, int32_t mode, const uint64_t* d_in, uint64_t* d_out, uint64_t count,
                         ps_stream s) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (!p) return fail(PS_ERR_INVALID_ARGUMENT, "puzzle is NULL");
    const int slots = slots_per_parent(p, mode);
    if (slots < 0) return fail(PS_ERR_INVALID_ARGUMENT, "mode %d not valid for this puzzle", mode);
    if (count == 0) return PS_OK;
    if (!d_in || !d_out) return fail(PS_ERR_INVALID_ARGUMENT, "NULL device buffer");
    if (count > (1ull << 40)) return fail(PS_ERR_INVALID_ARGUMENT, "count too large");
    ps_puzzle_info info;
    puzzle_query(p, &info);
    const int ops = info.operators_count;
    const psd::PuzzleParams prm = params_of(*p);
    hipStream_t st = stream_of(s);
    return dispatch<FAM_ALL>(*p, mode, [&](auto tag) -> int {
        using PZ = typename decltype(tag)::type;
        constexpr int TPB = 256;
        const uint64_t blocks = (count + TPB - 1) / TPB;
        hipLaunchKernelGGL((k_expand_slots<PZ, TPB>), dim3((uint32_t)blocks), dim3(TPB), 0, st, d_in, d_out, count,
                           ops, prm);
        PS_HIP_CHECK(hipGetLastError());
        return PS_OK;
    });
    PS_API_END
}

extern "C" int ps_pancake_xseg_postprocess(int32_t n, int32_t segment, uint32_t* d_idx, uint64_t count, ps_stream s) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (n < 13 || n > 16) return fail(PS_ERR_INVALID_ARGUMENT, "cross-segment post-process needs 13 <= n <= 16");
    if (segment < 0) return fail(PS_ERR_INVALID_ARGUMENT, "negative segment");
    if (count == 0) return PS_OK;
    if (!d_idx) return fail(PS_ERR_INVALID_ARGUMENT, "NULL device buffer");
    const uint32_t blocks = (uint32_t)((count + 255) / 256);
    hipStream_t st = stream_of(s);
    switch (n) {
    case 13: hipLaunchKernelGGL(k_pancake_xseg_post<13>, dim3(blocks), dim3(256), 0, st, d_idx, (uint32_t)segment, count); break;
    case 14: hipLaunchKernelGGL(k_pancake_xseg_post<14>, dim3(blocks), dim3(256), 0, st, d_idx, (uint32_t)segment, count); break;
    case 15: hipLaunchKernelGGL(k_pancake_xseg_post<15>, dim3(blocks), dim3(256), 0, st, d_idx, (uint32_t)segment, count); break;
    case 16: hipLaunchKernelGGL(k_pancake_xseg_post<16>, dim3(blocks), dim3(256), 0, st, d_idx, (uint32_t)segment, count); break;
    }
    PS_HIP_CHECK(hipGetLastError());
    return PS_OK;
    PS_API_END
}
