// Maps a runtime ps_puzzle (+ expand mode) to the compile-time device functor.
#pragma once

#include <cstdint>

#include "device_puzzles.h"
#include "ps_internal.h"

namespace ps {

template <class T>
struct Tag {
    using type = T;
};

enum : int { FAM_PANCAKE = 1, FAM_SLIDING = 2, FAM_PACKED = 4, FAM_HANOI = 8, FAM_ALL = 15 };

// Calls fn(Tag<Functor>{}) and returns its result, or fails with
// PS_ERR_UNSUPPORTED when the puzzle/mode has no functor in families FAM.
template <int FAM, class Fn>
int dispatch(const ps_puzzle& p, int mode, Fn&& fn) {
    using namespace psd;
    switch (p.kind) {
    case PS_PANCAKE:
        if constexpr ((FAM & FAM_PANCAKE) != 0) {
            if (mode == PS_EXPAND_FULL) {
                switch (p.size) {
                case 2: return fn(Tag<Pancake<2, 0>>{});
                case 3: return fn(Tag<Pancake<3, 0>>{});
                case 4: return fn(Tag<Pancake<4, 0>>{});
                case 5: return fn(Tag<Pancake<5, 0>>{});
                case 6: return fn(Tag<Pancake<6, 0>>{});
                case 7: return fn(Tag<Pancake<7, 0>>{});
                case 8: return fn(Tag<Pancake<8, 0>>{});
                case 9: return fn(Tag<Pancake<9, 0>>{});
                case 10: return fn(Tag<Pancake<10, 0>>{});
                case 11: return fn(Tag<Pancake<11, 0>>{});
                case 12: return fn(Tag<Pancake<12, 0>>{});
                case 13: return fn(Tag<Pancake<13, 0>>{});
                case 14: return fn(Tag<Pancake<14, 0>>{});
                case 15: return fn(Tag<Pancake<15, 0>>{});
                case 16: return fn(Tag<Pancake<16, 0>>{});
                }
            } else if (mode == PS_EXPAND_IN_SEGMENT) {
                switch (p.size) {
                case 13: return fn(Tag<Pancake<13, 1>>{});
                case 14: return fn(Tag<Pancake<14, 1>>{});
                case 15: return fn(Tag<Pancake<15, 1>>{});
                case 16: return fn(Tag<Pancake<16, 1>>{});
                }
            } else if (mode == PS_EXPAND_CROSS_SEGMENT) {
                switch (p.size) {
                case 13: return fn(Tag<Pancake<13, 2>>{});
                case 14: return fn(Tag<Pancake<14, 2>>{});
                case 15: return fn(Tag<Pancake<15, 2>>{});
                case 16: return fn(Tag<Pancake<16, 2>>{});
                }
            }
        }
        break;
    case PS_SLIDING:
        if constexpr ((FAM & FAM_SLIDING) != 0) {
            if (mode != PS_EXPAND_FULL || p.size < 1 || p.height < 1) break;
            switch (p.size * p.height) {
            case 2: return fn(Tag<Sliding<2>>{});
            case 3: return fn(Tag<Sliding<3>>{});
            case 4: return fn(Tag<Sliding<4>>{});
            case 5: return fn(Tag<Sliding<5>>{});
            case 6: return fn(Tag<Sliding<6>>{});
            case 7: return fn(Tag<Sliding<7>>{});
            case 8: return fn(Tag<Sliding<8>>{});
            case 9: return fn(Tag<Sliding<9>>{});
            case 10: return fn(Tag<Sliding<10>>{});
            case 11: return fn(Tag<Sliding<11>>{});
            case 12: return fn(Tag<Sliding<12>>{});
            case 13: return fn(Tag<Sliding<13>>{});
            case 14: return fn(Tag<Sliding<14>>{});
            case 15: return fn(Tag<Sliding<15>>{});
            case 16: return fn(Tag<Sliding<16>>{});
            }
        }
        break;
    case PS_SLIDING_PACKED:
        if constexpr ((FAM & FAM_PACKED) != 0) {
            if (mode != PS_EXPAND_FULL) break;
            // the instantiations of Puzzles/SlidingTilePuzzleOptimized.cpp:297-309
            const int key = p.size * 16 + p.height;
            switch (key) {
            case 2 * 16 + 2: return fn(Tag<Packed<2, 2>>{});
            case 3 * 16 + 2: return fn(Tag<Packed<3, 2>>{});
            case 4 * 16 + 2: return fn(Tag<Packed<4, 2>>{});
            case 5 * 16 + 2: return fn(Tag<Packed<5, 2>>{});
            case 6 * 16 + 2: return fn(Tag<Packed<6, 2>>{});
            case 7 * 16 + 2: return fn(Tag<Packed<7, 2>>{});
            case 8 * 16 + 2: return fn(Tag<Packed<8, 2>>{});
            case 3 * 16 + 3: return fn(Tag<Packed<3, 3>>{});
            case 4 * 16 + 3: return fn(Tag<Packed<4, 3>>{});
            case 5 * 16 + 3: return fn(Tag<Packed<5, 3>>{});
            case 4 * 16 + 4: return fn(Tag<Packed<4, 4>>{});
            }
        }
        break;
    case PS_HANOI4:
    case PS_HANOI4_OPT:
        if constexpr ((FAM & FAM_HANOI) != 0) {
            if (mode != PS_EXPAND_FULL) break;
            if (p.size >= 1 && p.size <= 16) return fn(Tag<Hanoi<uint32_t>>{});
            if (p.size > 16 && p.size <= 24) return fn(Tag<Hanoi<uint64_t>>{});
        }
        break;
    }
    return fail(PS_ERR_UNSUPPORTED, "no device expander for puzzle kind=%d size=%d height=%d mode=%d", p.kind,
                p.size, p.height, mode);
}

inline int family_of(const ps_puzzle& p) {
    switch (p.kind) {
    case PS_PANCAKE: return FAM_PANCAKE;
    case PS_SLIDING: return FAM_SLIDING;
    case PS_SLIDING_PACKED: return FAM_PACKED;
    case PS_HANOI4:
    case PS_HANOI4_OPT: return FAM_HANOI;
    }
    return 0;
}

inline psd::PuzzleParams params_of(const ps_puzzle& p) { return psd::PuzzleParams{p.kind, p.size, p.height, p.flags}; }

}  // namespace ps
