// Puzzles the multi-search kernels are compiled for: the serving mix
// (sliding tile 2x4/4x2, 3x3, 2x5/5x2, 3x4/4x3; pancake 8..12; 4-peg Hanoi up
// to 16 disks, with or without symmetry).  A search of another puzzle runs on
// its own ps_bfs instead.
#pragma once

#include "bfs_kernels.h"
#include "dispatch.h"

namespace ps {

enum : uint32_t {
    MK_PANCAKE8,
    MK_PANCAKE9,
    MK_PANCAKE10,
    MK_PANCAKE11,
    MK_PANCAKE12,
    MK_SLIDING8,
    MK_SLIDING9,
    MK_SLIDING10,
    MK_SLIDING12,
    MK_HANOI,
    MK_COUNT
};

// the kernel-table index of a puzzle, or -1
inline int multi_kind_of(const ps_puzzle& p) {
    switch (p.kind) {
    case PS_PANCAKE:
        if (p.flags & PS_FLAG_INV_INDEX) return -1;
        if (p.size >= 8 && p.size <= 12) return (int)MK_PANCAKE8 + (p.size - 8);
        return -1;
    case PS_SLIDING:
        switch (p.size * p.height) {
        case 8: return MK_SLIDING8;
        case 9: return MK_SLIDING9;
        case 10: return MK_SLIDING10;
        case 12: return MK_SLIDING12;
        }
        return -1;
    case PS_HANOI4:
    case PS_HANOI4_OPT:
        return (p.size >= 1 && p.size <= 16) ? (int)MK_HANOI : -1;
    }
    return -1;
}

struct MultiKinds {
    template <int TPB>
    __device__ __forceinline__ static void run(uint32_t kind, const psk::ExpandArgs& a, uint64_t tile,
                                               const psk::ExpandLds& lds, uint32_t* s_local, psk::ExpandTally& t) {
        using namespace psd;
        switch (kind) {
        case MK_PANCAKE8: psk::expand_tile<Pancake<8, 0>, TPB>(a, tile, lds, s_local, t); break;
        case MK_PANCAKE9: psk::expand_tile<Pancake<9, 0>, TPB>(a, tile, lds, s_local, t); break;
        case MK_PANCAKE10: psk::expand_tile<Pancake<10, 0>, TPB>(a, tile, lds, s_local, t); break;
        case MK_PANCAKE11: psk::expand_tile<Pancake<11, 0>, TPB>(a, tile, lds, s_local, t); break;
        case MK_PANCAKE12: psk::expand_tile<Pancake<12, 0>, TPB>(a, tile, lds, s_local, t); break;
        case MK_SLIDING8: psk::expand_tile<Sliding<8>, TPB>(a, tile, lds, s_local, t); break;
        case MK_SLIDING9: psk::expand_tile<Sliding<9>, TPB>(a, tile, lds, s_local, t); break;
        case MK_SLIDING10: psk::expand_tile<Sliding<10>, TPB>(a, tile, lds, s_local, t); break;
        case MK_SLIDING12: psk::expand_tile<Sliding<12>, TPB>(a, tile, lds, s_local, t); break;
        default: psk::expand_tile<Hanoi<uint32_t>, TPB>(a, tile, lds, s_local, t); break;
        }
    }
};

}  // namespace ps
