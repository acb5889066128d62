// Instantiates the device BFS and segment-step kernels for the pancake puzzle
// family, and the cross-tile list kernels of 10..12 pancakes.
#include "bfs_launch.h"
#include "bfs_pancake_lists.h"
#include "fs_launch.h"
#include "seg3_launch.h"

namespace ps {
PS_DEFINE_BFS_FAMILY(pancake, FAM_PANCAKE)
PS_DEFINE_FS_FAMILY(pancake, FAM_PANCAKE)
PS_DEFINE_S3_FAMILY(pancake, FAM_PANCAKE)

namespace {

template <class K, class... Args>
int launch_pkl(K kern, const LaunchCfg& c, int* occ, const Args&... args) {
    if (occ) return launch_or_query(kern, c, occ, nullptr);
    hipLaunchKernelGGL(kern, dim3(c.grid), dim3(c.tpb), c.lds, c.stream, args...);
    PS_HIP_CHECK(hipGetLastError());
    return PS_OK;
}

bool pkl_size(const ps_puzzle& p) {
    return p.kind == PS_PANCAKE && p.size >= 10 && p.size <= 12;
}

}  // namespace

int pkl_expand(const ps_puzzle& p, const psk::ExpandArgs& a, const psk::ListArgs& x, const LaunchCfg& c, int* occ) {
    if (!pkl_size(p)) return fail(PS_ERR_UNSUPPORTED, "cross-tile lists cover 10..12 pancakes");
    switch (p.size * (c.tpb >= 1000 ? 10000 : 1000) + c.tpb) {
    case 10256: return launch_pkl(psk::k_pkl_expand<10, 256>, c, occ, a, x);
    case 10512: return launch_pkl(psk::k_pkl_expand<10, 512>, c, occ, a, x);
    case 11256: return launch_pkl(psk::k_pkl_expand<11, 256>, c, occ, a, x);
    case 11512: return launch_pkl(psk::k_pkl_expand<11, 512>, c, occ, a, x);
    case 12256: return launch_pkl(psk::k_pkl_expand<12, 256>, c, occ, a, x);
    case 12512: return launch_pkl(psk::k_pkl_expand<12, 512>, c, occ, a, x);
    case 101024: return launch_pkl(psk::k_pkl_expand<10, 1024>, c, occ, a, x);
    case 111024: return launch_pkl(psk::k_pkl_expand<11, 1024>, c, occ, a, x);
    case 121024: return launch_pkl(psk::k_pkl_expand<12, 1024>, c, occ, a, x);
    }
    return fail(PS_ERR_INVALID_ARGUMENT, "block_threads %d not in {256, 512, 1024}", c.tpb);
}

int pkl_apply(const ps_puzzle& p, const psk::ApplyArgs& a, const LaunchCfg& c, int* occ) {
    if (!pkl_size(p)) return fail(PS_ERR_UNSUPPORTED, "cross-tile lists cover 10..12 pancakes");
    switch (p.size * 10000 + c.tpb) {
    case 100512: return launch_pkl(psk::k_pkl_apply<10, 512>, c, occ, a);
    case 110512: return launch_pkl(psk::k_pkl_apply<11, 512>, c, occ, a);
    case 120512: return launch_pkl(psk::k_pkl_apply<12, 512>, c, occ, a);
    case 101024: return launch_pkl(psk::k_pkl_apply<10, 1024>, c, occ, a);
    case 111024: return launch_pkl(psk::k_pkl_apply<11, 1024>, c, occ, a);
    case 121024: return launch_pkl(psk::k_pkl_apply<12, 1024>, c, occ, a);
    }
    if (c.tpb != 512 && c.tpb != 1024) return fail(PS_ERR_INVALID_ARGUMENT, "the list apply kernel runs 512 or 1024 threads");
    return fail(PS_ERR_UNSUPPORTED, "cross-tile lists cover 10..12 pancakes");
}

int pkl_step(const ps_puzzle& p, const psk::StepArgs& a, const LaunchCfg& c, int* occ) {
    if (!pkl_size(p)) return fail(PS_ERR_UNSUPPORTED, "cross-tile lists cover 10..12 pancakes");
    const int key = p.size * 10000 + c.tpb;
    if (a.role == 1) {
        switch (key) {
        case 100512: return launch_pkl(psk::k_pkl_step<10, 512, 1>, c, occ, a);
        case 101024: return launch_pkl(psk::k_pkl_step<10, 1024, 1>, c, occ, a);
        case 110512: return launch_pkl(psk::k_pkl_step<11, 512, 1>, c, occ, a);
        case 111024: return launch_pkl(psk::k_pkl_step<11, 1024, 1>, c, occ, a);
        case 120512: return launch_pkl(psk::k_pkl_step<12, 512, 1>, c, occ, a);
        case 121024: return launch_pkl(psk::k_pkl_step<12, 1024, 1>, c, occ, a);
        }
        return fail(PS_ERR_INVALID_ARGUMENT, "split list layers run 512 or 1024 threads, not %d", c.tpb);
    }
    switch (key) {
    case 100256: return launch_pkl(psk::k_pkl_step<10, 256>, c, occ, a);
    case 100512: return launch_pkl(psk::k_pkl_step<10, 512>, c, occ, a);
    case 101024: return launch_pkl(psk::k_pkl_step<10, 1024>, c, occ, a);
    case 110256: return launch_pkl(psk::k_pkl_step<11, 256>, c, occ, a);
    case 110512: return launch_pkl(psk::k_pkl_step<11, 512>, c, occ, a);
    case 111024: return launch_pkl(psk::k_pkl_step<11, 1024>, c, occ, a);
    case 120256: return launch_pkl(psk::k_pkl_step<12, 256>, c, occ, a);
    case 120512: return launch_pkl(psk::k_pkl_step<12, 512>, c, occ, a);
    case 121024: return launch_pkl(psk::k_pkl_step<12, 1024>, c, occ, a);
    }
    return fail(PS_ERR_INVALID_ARGUMENT, "block_threads %d not in {256, 512, 1024}", c.tpb);
}

int pkl_pull(const ps_puzzle& p, const psk::PullArgs& a, const LaunchCfg& c, int* occ) {
    if (!pkl_size(p)) return fail(PS_ERR_UNSUPPORTED, "cross-tile lists cover 10..12 pancakes");
    switch (p.size * 10000 + c.tpb) {
    case 100512: return launch_pkl(psk::k_pkl_pull<10, 512>, c, occ, a);
    case 101024: return launch_pkl(psk::k_pkl_pull<10, 1024>, c, occ, a);
    case 110512: return launch_pkl(psk::k_pkl_pull<11, 512>, c, occ, a);
    case 111024: return launch_pkl(psk::k_pkl_pull<11, 1024>, c, occ, a);
    case 120512: return launch_pkl(psk::k_pkl_pull<12, 512>, c, occ, a);
    case 121024: return launch_pkl(psk::k_pkl_pull<12, 1024>, c, occ, a);
    }
    return fail(PS_ERR_INVALID_ARGUMENT, "the bottom-up list kernel runs 512 or 1024 threads, not %d", c.tpb);
}

size_t pkl_pull_lds_bytes(const ps_puzzle& p, int tpb) {
    if (!pkl_size(p)) return 0;
    switch (p.size) {
    case 10: return psk::pkl_pull_lds<10>(tpb);
    case 11: return psk::pkl_pull_lds<11>(tpb);
    case 12: return psk::pkl_pull_lds<12>(tpb);
    }
    return 0;
}

int pkl_layer_end(const psk::LayerEndArgs& a, const psk::StepArgs& s, int advance, hipStream_t stream) {
    hipLaunchKernelGGL(psk::k_pkl_layer_end<1024>, dim3(1), dim3(1024), 0, stream, a, s, advance);
    PS_HIP_CHECK(hipGetLastError());
    return PS_OK;
}

size_t pkl_step_lds_bytes(const ps_puzzle& p, int tpb) {
    if (!pkl_size(p)) return 0;
    switch (p.size) {
    case 10: return psk::pkl_step_lds<10>(tpb);
    case 11: return psk::pkl_step_lds<11>(tpb);
    case 12: return psk::pkl_step_lds<12>(tpb);
    }
    return 0;
}

size_t pkl_expand_lds_bytes(const ps_puzzle& p, int tpb, uint32_t nseg) {
    if (!pkl_size(p)) return 0;
    switch (p.size) {
    case 10: return psk::pkl_expand_lds<10>(tpb, nseg);
    case 11: return psk::pkl_expand_lds<11>(tpb, nseg);
    case 12: return psk::pkl_expand_lds<12>(tpb, nseg);
    }
    return 0;
}

size_t pkl_apply_lds_bytes(const ps_puzzle& p, int tpb) {
    if (!pkl_size(p)) return 0;
    switch (p.size) {
    case 10: return psk::pkl_apply_lds<10>(tpb);
    case 11: return psk::pkl_apply_lds<11>(tpb);
    case 12: return psk::pkl_apply_lds<12>(tpb);
    }
    return 0;
}

}  // namespace ps
