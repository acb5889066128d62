// Launchers for the templated BFS kernels; each puzzle family's instantiations
// live in their own translation unit (bfs_pancake.hip, bfs_sliding.hip,
// bfs_hanoi.hip) to keep compile times down.
#pragma once

#include "bfs_kernels.h"
#include "dispatch.h"

namespace ps {

struct LaunchCfg {
    int tpb;
    uint32_t grid;
    size_t lds;
    hipStream_t stream;
};

// occ != NULL: only report resident blocks per CU for this configuration (and
// raise the kernel's dynamic-LDS limit); otherwise launch.
#define PS_DECLARE_BFS_FAMILY(NAME)                                                                   \
    int bfs3_expand_##NAME(const ps_puzzle& p, const psk::ExpandArgs& a, const LaunchCfg& c, int* occ); \
    int bfs3_tile_step_##NAME(const ps_puzzle& p, const psk::ExpandArgs& a, const psk::TileStepArgs& t,  \
                              const LaunchCfg& c, int* occ);                                             \
    int bfs2_layer_##NAME(const ps_puzzle& p, const psk::Bfs2Args& a, const LaunchCfg& c, int* occ);
PS_DECLARE_BFS_FAMILY(pancake)
PS_DECLARE_BFS_FAMILY(sliding)
PS_DECLARE_BFS_FAMILY(packed)
PS_DECLARE_BFS_FAMILY(hanoi)

// Cross-tile list kernels of 10..12 pancakes (bfs_pancake_lists.h), defined in
// bfs_pancake.hip.  occ != NULL: report resident blocks per CU only.
int pkl_expand(const ps_puzzle& p, const psk::ExpandArgs& a, const psk::ListArgs& x, const LaunchCfg& c, int* occ);
int pkl_apply(const ps_puzzle& p, const psk::ApplyArgs& a, const LaunchCfg& c, int* occ);
int pkl_step(const ps_puzzle& p, const psk::StepArgs& a, const LaunchCfg& c, int* occ);
// bottom-up expansion of a split layer (k_pkl_pull; it exits at once unless the
// finalize launch before it chose that direction)
int pkl_pull(const ps_puzzle& p, const psk::PullArgs& a, const LaunchCfg& c, int* occ);
size_t pkl_pull_lds_bytes(const ps_puzzle& p, int tpb);
// bookkeeping between two k_pkl_step launches (advance: also the layer end)
int pkl_layer_end(const psk::LayerEndArgs& a, const psk::StepArgs& s, int advance, hipStream_t stream);
// LDS bytes of the kernels; 0 when the puzzle has no list kernels
size_t pkl_expand_lds_bytes(const ps_puzzle& p, int tpb, uint32_t nseg);
size_t pkl_apply_lds_bytes(const ps_puzzle& p, int tpb);
size_t pkl_step_lds_bytes(const ps_puzzle& p, int tpb);

inline int bfs3_expand(const ps_puzzle& p, const psk::ExpandArgs& a, const LaunchCfg& c, int* occ) {
    switch (family_of(p)) {
    case FAM_PANCAKE: return bfs3_expand_pancake(p, a, c, occ);
    case FAM_SLIDING: return bfs3_expand_sliding(p, a, c, occ);
    case FAM_PACKED: return bfs3_expand_packed(p, a, c, occ);
    case FAM_HANOI: return bfs3_expand_hanoi(p, a, c, occ);
    }
    return fail(PS_ERR_UNSUPPORTED, "unknown puzzle family");
}

inline int bfs3_tile_step(const ps_puzzle& p, const psk::ExpandArgs& a, const psk::TileStepArgs& t,
                          const LaunchCfg& c, int* occ) {
    switch (family_of(p)) {
    case FAM_PANCAKE: return bfs3_tile_step_pancake(p, a, t, c, occ);
    case FAM_SLIDING: return bfs3_tile_step_sliding(p, a, t, c, occ);
    case FAM_PACKED: return bfs3_tile_step_packed(p, a, t, c, occ);
    case FAM_HANOI: return bfs3_tile_step_hanoi(p, a, t, c, occ);
    }
    return fail(PS_ERR_UNSUPPORTED, "unknown puzzle family");
}

inline int bfs2_layer(const ps_puzzle& p, const psk::Bfs2Args& a, const LaunchCfg& c, int* occ) {
    switch (family_of(p)) {
    case FAM_PANCAKE: return bfs2_layer_pancake(p, a, c, occ);
    case FAM_SLIDING: return bfs2_layer_sliding(p, a, c, occ);
    case FAM_PACKED: return bfs2_layer_packed(p, a, c, occ);
    case FAM_HANOI: return bfs2_layer_hanoi(p, a, c, occ);
    }
    return fail(PS_ERR_UNSUPPORTED, "unknown puzzle family");
}

template <class K>
int launch_or_query(K kern, const LaunchCfg& c, int* occ, const void* args_dummy) {
    (void)args_dummy;
    if (occ) {
        if (c.lds > 65536)
            PS_HIP_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)c.lds));
        PS_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, kern, c.tpb, c.lds));
        return PS_OK;
    }
    return PS_OK;
}

#define PS_DEFINE_BFS_FAMILY(NAME, FAMV)                                                                        \
    int bfs3_expand_##NAME(const ps_puzzle& p, const psk::ExpandArgs& a, const LaunchCfg& c, int* occ) {       \
        return dispatch<FAMV>(p, PS_EXPAND_FULL, [&](auto tag) -> int {                                         \
            using PZ = typename decltype(tag)::type;                                                            \
            auto go = [&](auto kern) -> int {                                                                   \
                if (occ) return launch_or_query(kern, c, occ, nullptr);                                         \
                hipLaunchKernelGGL(kern, dim3(c.grid), dim3(c.tpb), c.lds, c.stream, a);                        \
                PS_HIP_CHECK(hipGetLastError());                                                                \
                return PS_OK;                                                                                   \
            };                                                                                                  \
            switch (c.tpb) {                                                                                    \
            case 256: return go(psk::k_bfs3_expand<PZ, 256>);                                                   \
            case 512: return go(psk::k_bfs3_expand<PZ, 512>);                                                   \
            }                                                                                                   \
            return fail(PS_ERR_INVALID_ARGUMENT, "block_threads %d not in {256, 512}", c.tpb);                 \
        });                                                                                                     \
    }                                                                                                           \
    int bfs3_tile_step_##NAME(const ps_puzzle& p, const psk::ExpandArgs& a, const psk::TileStepArgs& t,        \
                              const LaunchCfg& c, int* occ) {                                                   \
        return dispatch<FAMV>(p, PS_EXPAND_FULL, [&](auto tag) -> int {                                         \
            using PZ = typename decltype(tag)::type;                                                            \
            auto go = [&](auto kern) -> int {                                                                   \
                if (occ) return launch_or_query(kern, c, occ, nullptr);                                         \
                hipLaunchKernelGGL(kern, dim3(c.grid), dim3(c.tpb), c.lds, c.stream, a, t);                     \
                PS_HIP_CHECK(hipGetLastError());                                                                \
                return PS_OK;                                                                                   \
            };                                                                                                  \
            switch (c.tpb) {                                                                                    \
            case 256: return go(psk::k_bfs3_tile_step<PZ, 256>);                                                \
            case 512: return go(psk::k_bfs3_tile_step<PZ, 512>);                                                \
            }                                                                                                   \
            return fail(PS_ERR_INVALID_ARGUMENT, "block_threads %d not in {256, 512}", c.tpb);                 \
        });                                                                                                     \
    }                                                                                                           \
    int bfs2_layer_##NAME(const ps_puzzle& p, const psk::Bfs2Args& a, const LaunchCfg& c, int* occ) {          \
        return dispatch<FAMV>(p, PS_EXPAND_FULL, [&](auto tag) -> int {                                         \
            using PZ = typename decltype(tag)::type;                                                            \
            auto kern = psk::k_bfs2_layer<PZ, 256>;                                                             \
            if (c.tpb != 256) return fail(PS_ERR_INVALID_ARGUMENT, "two-bit search uses 256 threads per block"); \
            if (occ) return launch_or_query(kern, c, occ, nullptr);                                             \
            hipLaunchKernelGGL(kern, dim3(c.grid), dim3(c.tpb), c.lds, c.stream, a);                            \
            PS_HIP_CHECK(hipGetLastError());                                                                    \
            return PS_OK;                                                                                       \
        });                                                                                                     \
    }

}  // namespace ps
