// Launchers for the templated segment-step kernels; each puzzle family's
// instantiations live in that family's translation unit (bfs_<family>.hip).
// Only puzzles with at most 4 moves have a 4-bit op-flag array
// (Common/DiskBasedOptFrontierSearch.cpp:257-260).
#pragma once

#include "dispatch.h"
#include "fs_kernels.h"

namespace ps {

constexpr int FS_TPB = 256;

#define PS_DECLARE_FS_FAMILY(NAME)                                                                       \
    int fs_bin_##NAME(const ps_puzzle& p, const psk::FsBinArgs& a, int place, uint32_t grid, hipStream_t s); \
    int fs_cross_##NAME(const ps_puzzle& p, const psk::FsCrossArgs& a, uint32_t grid, hipStream_t s, int* slots);
PS_DECLARE_FS_FAMILY(pancake)
PS_DECLARE_FS_FAMILY(sliding)
PS_DECLARE_FS_FAMILY(packed)
PS_DECLARE_FS_FAMILY(hanoi)

#define PS_DEFINE_FS_FAMILY(NAME, FAMV)                                                                       \
    int fs_bin_##NAME(const ps_puzzle& p, const psk::FsBinArgs& a, int place, uint32_t grid, hipStream_t s) { \
        return dispatch<FAMV>(p, PS_EXPAND_FULL, [&](auto tag) -> int {                                     \
            using PZ = typename decltype(tag)::type;                                                        \
            if constexpr (PZ::OPS <= 4) {                                                                   \
                const size_t lds = (size_t)a.nregions * 4;                                                  \
                if (place)                                                                                  \
                    hipLaunchKernelGGL((psk::k_fs_bin<PZ, psk::FS_BIN_TPB, true>), dim3(grid),                \
                                       dim3(psk::FS_BIN_TPB), lds, s, a);                                   \
                else                                                                                        \
                    hipLaunchKernelGGL((psk::k_fs_bin<PZ, psk::FS_BIN_TPB, false>), dim3(grid),               \
                                       dim3(psk::FS_BIN_TPB), lds, s, a);                                   \
                PS_HIP_CHECK(hipGetLastError());                                                            \
                return PS_OK;                                                                               \
            } else {                                                                                        \
                return fail(PS_ERR_UNSUPPORTED, "the segment step needs at most 4 moves (puzzle has %d)",   \
                            PZ::OPS);                                                                       \
            }                                                                                               \
        });                                                                                                 \
    }                                                                                                       \
    int fs_cross_##NAME(const ps_puzzle& p, const psk::FsCrossArgs& a, uint32_t grid, hipStream_t s,       \
                        int* slots) {                                                                       \
        return dispatch<FAMV>(p, PS_EXPAND_FULL, [&](auto tag) -> int {                                     \
            using PZ = typename decltype(tag)::type;                                                        \
            if constexpr (PZ::OPS <= 4) {                                                                   \
                if (slots) {                                                                                \
                    *slots = PZ::SLOTS;                                                                     \
                    return PS_OK;                                                                           \
                }                                                                                           \
                hipLaunchKernelGGL((psk::k_fs_cross<PZ, FS_TPB>), dim3(grid), dim3(FS_TPB), 0, s, a);       \
                PS_HIP_CHECK(hipGetLastError());                                                            \
                return PS_OK;                                                                               \
            } else {                                                                                        \
                return fail(PS_ERR_UNSUPPORTED, "the segment step needs at most 4 moves (puzzle has %d)",   \
                            PZ::OPS);                                                                       \
            }                                                                                               \
        });                                                                                                 \
    }

// the column pass over the (workgroup, region) table of the bin passes
// (k_fs_bin_columns): rcount[r] = marks of region r, and each entry replaced
// by the workgroup's first position in its region
int fs_bin_columns(uint32_t* table, uint32_t rows, uint32_t nregions, uint32_t* rcount, hipStream_t s);

// place = 0: count the marks per (workgroup, region); 1: write them to their regions' ranges
inline int fs_bin(const ps_puzzle& p, const psk::FsBinArgs& a, int place, uint32_t grid, hipStream_t s) {
    switch (family_of(p)) {
    case FAM_PANCAKE: return fs_bin_pancake(p, a, place, grid, s);
    case FAM_SLIDING: return fs_bin_sliding(p, a, place, grid, s);
    case FAM_PACKED: return fs_bin_packed(p, a, place, grid, s);
    case FAM_HANOI: return fs_bin_hanoi(p, a, place, grid, s);
    }
    return fail(PS_ERR_UNSUPPORTED, "unknown puzzle family");
}

// slots != NULL: only report the slots per parent
inline int fs_cross(const ps_puzzle& p, const psk::FsCrossArgs& a, uint32_t grid, hipStream_t s, int* slots) {
    switch (family_of(p)) {
    case FAM_PANCAKE: return fs_cross_pancake(p, a, grid, s, slots);
    case FAM_SLIDING: return fs_cross_sliding(p, a, grid, s, slots);
    case FAM_PACKED: return fs_cross_packed(p, a, grid, s, slots);
    case FAM_HANOI: return fs_cross_hanoi(p, a, grid, s, slots);
    }
    return fail(PS_ERR_UNSUPPORTED, "unknown puzzle family");
}

}  // namespace ps
