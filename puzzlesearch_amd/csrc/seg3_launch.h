// Launchers for the templated kernels of the three-bit segment step; each
// puzzle family's instantiations live in that family's translation unit
// (bfs_<family>.hip).  `mode` is the expansion of the reference's ExpandHint:
// PS_EXPAND_IN_SEGMENT / PS_EXPAND_CROSS_SEGMENT for more than 12 pancakes in
// 29-bit segments, PS_EXPAND_FULL otherwise.
#pragma once

#include "dispatch.h"
#include "seg3_kernels.h"

namespace ps {

constexpr int S3_TPB = 256;

#define PS_DECLARE_S3_FAMILY(NAME)                                                                            \
    int s3_mark_in_##NAME(const ps_puzzle& p, int mode, const psk::S3MarkArgs& a, uint32_t grid, hipStream_t s); \
    int s3_cross_##NAME(const ps_puzzle& p, int mode, const psk::S3CrossArgs& a, uint32_t grid, hipStream_t s,  \
                        int* slots);
PS_DECLARE_S3_FAMILY(pancake)
PS_DECLARE_S3_FAMILY(sliding)
PS_DECLARE_S3_FAMILY(packed)
PS_DECLARE_S3_FAMILY(hanoi)

#define PS_DEFINE_S3_FAMILY(NAME, FAMV)                                                                             \
    int s3_mark_in_##NAME(const ps_puzzle& p, int mode, const psk::S3MarkArgs& a, uint32_t grid, hipStream_t s) { \
        return dispatch<FAMV>(p, mode, [&](auto tag) -> int {                                                     \
            using PZ = typename decltype(tag)::type;                                                              \
            hipLaunchKernelGGL((psk::k_s3_mark_in<PZ, S3_TPB>), dim3(grid), dim3(S3_TPB), 0, s, a);               \
            PS_HIP_CHECK(hipGetLastError());                                                                      \
            return PS_OK;                                                                                         \
        });                                                                                                       \
    }                                                                                                             \
    int s3_cross_##NAME(const ps_puzzle& p, int mode, const psk::S3CrossArgs& a, uint32_t grid, hipStream_t s,    \
                        int* slots) {                                                                             \
        return dispatch<FAMV>(p, mode, [&](auto tag) -> int {                                                     \
            using PZ = typename decltype(tag)::type;                                                              \
            if (slots) {                                                                                          \
                *slots = PZ::SLOTS;                                                                               \
                return PS_OK;                                                                                     \
            }                                                                                                     \
            hipLaunchKernelGGL((psk::k_s3_cross<PZ, S3_TPB>), dim3(grid), dim3(S3_TPB), 0, s, a);                 \
            PS_HIP_CHECK(hipGetLastError());                                                                      \
            return PS_OK;                                                                                         \
        });                                                                                                       \
    }

inline int s3_mark_in(const ps_puzzle& p, int mode, const psk::S3MarkArgs& a, uint32_t grid, hipStream_t s) {
    switch (family_of(p)) {
    case FAM_PANCAKE: return s3_mark_in_pancake(p, mode, a, grid, s);
    case FAM_SLIDING: return s3_mark_in_sliding(p, mode, a, grid, s);
    case FAM_PACKED: return s3_mark_in_packed(p, mode, a, grid, s);
    case FAM_HANOI: return s3_mark_in_hanoi(p, mode, a, grid, s);
    }
    return fail(PS_ERR_UNSUPPORTED, "unknown puzzle family");
}

// slots != NULL: only report the slots per parent of this mode
inline int s3_cross(const ps_puzzle& p, int mode, const psk::S3CrossArgs& a, uint32_t grid, hipStream_t s,
                    int* slots) {
    switch (family_of(p)) {
    case FAM_PANCAKE: return s3_cross_pancake(p, mode, a, grid, s, slots);
    case FAM_SLIDING: return s3_cross_sliding(p, mode, a, grid, s, slots);
    case FAM_PACKED: return s3_cross_packed(p, mode, a, grid, s, slots);
    case FAM_HANOI: return s3_cross_hanoi(p, mode, a, grid, s, slots);
    }
    return fail(PS_ERR_UNSUPPORTED, "unknown puzzle family");
}

}  // namespace ps
