// Device move generators for the puzzles of the reference's GPU plugins.
//
// Each functor has
//   static constexpr int SLOTS;                       // children slots per parent
//   template <class F> static void expand(index, op_bits, prm, F&& f);
// and calls f(slot, child_index, op_label) for every child that exists and whose
// operator bit is not set in op_bits, in slot order.  The slot/label layout is
// the reference kernels' layout; the arithmetic is this library's own:
//
//  * pancake: a flip of the last k elements changes only the last k Lehmer
//    digits, so child = rank - (rank mod k!) + rank(reversed suffix pattern).
//    All k are produced from one right-to-left sweep over the digits; pairwise
//    order comes from the digits themselves (a[s] < a[r]  <=>  c(s+1,r) + d[r]
//    >= d[s]), so the permutation is never materialised.
//  * sliding tile (lexicographic rank): a horizontal blank move changes two
//    adjacent digits (closed form); vertical moves re-rank the child.
//  * sliding tile (packed 32-bit segment encoding): the reference's
//    pack/unpack/rotate steps with every loop unrolled over the board size.
//  * 4-peg Hanoi: base-4 index -> four peg bit-planes; moving the top disk t
//    from peg p to q is index ^ ((p ^ q) << t); symmetry relabels pegs 1..3.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace psd {

constexpr uint64_t kInvalid = ~0ull;

struct PuzzleParams {
    int32_t kind, size, height, flags;
};

constexpr uint64_t kFact[21] = {1ull,
                                1ull,
                                2ull,
                                6ull,
                                24ull,
                                120ull,
                                720ull,
                                5040ull,
                                40320ull,
                                362880ull,
                                3628800ull,
                                39916800ull,
                                479001600ull,
                                6227020800ull,
                                87178291200ull,
                                1307674368000ull,
                                20922789888000ull,
                                355687428096000ull,
                                6402373705728000ull,
                                121645100408832000ull,
                                2432902008176640000ull};

// c * f for a small count c (< 2^8) and a constant factorial f: 24-bit multiply
// (full rate) when f fits in 24 bits.
__device__ __forceinline__ uint32_t mul_small(uint32_t c, uint32_t f) {
    return (f < (1u << 24)) ? __umul24(c, f) : c * f;
}

// ------------------------------------------------------------------ pancake

// Lehmer digits of the PancakeOptimized rank (Puzzles/PancakeOptimized.cpp:54-77):
// n <= 12 plain lexicographic rank; n > 12 segment = index >> 29 holds digits
// 0..n-13 (radices n..13), the low 29 bits digits n-12..n-1 (radices 12..1).
template <int N>
__device__ __forceinline__ void pancake_digits(uint64_t index, uint32_t (&d)[N]) {
    if constexpr (N <= 12) {
        uint32_t r = (uint32_t)index;
#pragma unroll
        for (int i = N - 1; i >= 0; --i) {
            const uint32_t m = (uint32_t)(N - i);
            d[i] = r % m;
            r /= m;
        }
    } else {
        uint32_t seg = (uint32_t)(index >> 29);
        uint32_t low = (uint32_t)(index & ((1u << 29) - 1));
#pragma unroll
        for (int i = N - 1; i >= N - 12; --i) {
            const uint32_t m = (uint32_t)(N - i);
            d[i] = low % m;
            low /= m;
        }
#pragma unroll
        for (int i = N - 13; i >= 0; --i) {
            const uint32_t m = (uint32_t)(N - i);
            d[i] = seg % m;
            seg /= m;
        }
    }
}

// Index from Lehmer digits in the same encoding.
template <int N>
__device__ __forceinline__ uint64_t pancake_index_from_digits(const uint32_t (&d)[N]) {
    if constexpr (N <= 12) {
        uint32_t r = 0;
#pragma unroll
        for (int i = 0; i < N; ++i) r += mul_small(d[i], (uint32_t)kFact[N - 1 - i]);
        return r;
    } else {
        uint32_t low = 0;
#pragma unroll
        for (int i = N - 12; i < N; ++i) low += mul_small(d[i], (uint32_t)kFact[N - 1 - i]);
        uint32_t seg = 0;
#pragma unroll
        for (int i = 0; i < N - 12; ++i) seg = seg * (uint32_t)(N - i) + d[i];
        return ((uint64_t)seg << 29) | low;
    }
}

// Digits <-> permutation (slow path, only for the inverse-index variant).
template <int N>
__device__ __forceinline__ void digits_to_perm(const uint32_t (&d)[N], uint32_t (&a)[N]) {
    uint64_t tiles = 0xFEDCBA9876543210ull;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const uint32_t sh = d[i] * 4;
        a[i] = (uint32_t)((tiles >> sh) & 15);
        const uint64_t lo = tiles & ((1ull << sh) - 1ull);
        const uint64_t hi = (tiles >> sh >> 4) << sh;
        tiles = hi | lo;
    }
}

template <int N>
__device__ __forceinline__ void perm_to_digits(const uint32_t (&a)[N], uint32_t (&d)[N]) {
    uint32_t seen = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        d[i] = a[i] - __popc(seen & ((1u << a[i]) - 1u));
        seen |= 1u << a[i];
    }
}

template <int N, int MODE>
struct Pancake {
    // MODE 0: all flips (ops 0..N-2); 1: in-segment ops 0..10; 2: cross-segment ops 11..N-2
    static constexpr int OPS = N - 1;
    static constexpr int OP_LO = (MODE == 2) ? 11 : 0;
    static constexpr int OP_HI = (MODE == 1) ? 11 : N - 1;
    static constexpr int SLOTS = OP_HI - OP_LO;

    template <class F>
    __device__ __forceinline__ static void expand(uint64_t index, uint32_t op_bits, const PuzzleParams& prm,
                                                  F&& f) {
        uint32_t d[N];
        pancake_digits<N>(index, d);
        const bool inv = (MODE == 2) && (prm.flags & 2) && N > 12;
        // cc[r] = #{q in [s, r): a[q] < a[r]} for the current suffix start s
        uint32_t cc[N];
#pragma unroll
        for (int r = 0; r < N; ++r) cc[r] = 0;
        // low-part value of the parent and its suffix digit sums
        const uint32_t low = (N <= 12) ? (uint32_t)index : (uint32_t)(index & ((1u << 29) - 1));
        const uint64_t high = (N <= 12) ? 0ull : (index & ~((1ull << 29) - 1ull));
        uint32_t suffix = 0;  // sum_{i >= s} d[i] * (N-1-i)!   (valid while N-s <= 12)
#pragma unroll
        for (int s = N - 2; s >= 0; --s) {
            const int k = N - s;  // flip length
            const int op = k - 2;
            uint32_t rho = 0;  // rank of the reversed suffix pattern: sum_r c(s,r) (r-s)!
#pragma unroll
            for (int r = s + 1; r < N; ++r) {
                const uint32_t lt = (cc[r] + d[r] >= d[s]) ? 1u : 0u;
                cc[r] += lt;
                if (k <= 12) rho += mul_small(cc[r], (uint32_t)kFact[r - s]);
            }
            if (k <= 12) suffix += mul_small(d[s], (uint32_t)kFact[N - 1 - s]);
            if (op < OP_LO || op >= OP_HI) continue;
            if (op_bits & (1u << op)) continue;
            uint64_t child;
            if (!inv && k <= 12) {
                child = high | (uint64_t)(low - suffix + rho);
            } else {
                // child digits: prefix unchanged; position P >= s holds c(s, s + N-1-P)
                uint32_t cd[N];
#pragma unroll
                for (int P = 0; P < N; ++P) cd[P] = (P < s) ? d[P] : cc[s + N - 1 - P];
                if (inv) {
                    uint32_t a[N];
                    digits_to_perm<N>(cd, a);
#pragma unroll
                    for (int i = 0; i < 6; ++i) {  // reverse the last 12 elements
                        const uint32_t t = a[N - 12 + i];
                        a[N - 12 + i] = a[N - 1 - i];
                        a[N - 1 - i] = t;
                    }
                    perm_to_digits<N>(a, cd);
                }
                child = pancake_index_from_digits<N>(cd);
            }
            f(op - OP_LO, child, (uint32_t)op);
        }
    }
};

// ---------------------------------------------------------- sliding (simple)

constexpr int OP_UP = 0, OP_LEFT = 1, OP_RIGHT = 2, OP_DOWN = 3;

// SlidingTilePuzzleGpu: lexicographic rank over N! (Puzzles/PermutationMap.cpp).
template <int N>
struct Sliding {
    static constexpr int OPS = 4;
    static constexpr int SLOTS = 4;
    using R = typename std::conditional<(N <= 12), uint32_t, uint64_t>::type;

    __device__ __forceinline__ static void digits(R r, uint32_t (&d)[N]) {
#pragma unroll
        for (int i = N - 1; i >= 0; --i) {
            const R m = (R)(N - i);
            d[i] = (uint32_t)(r % m);
            r /= m;
        }
    }

    // A vertical move swaps the blank with the tile `width` positions away.
    // Only the Lehmer digits of the two positions p < q and of the width-1
    // positions between them change: with t the moving tile and `less` the
    // tiles between them smaller than t,
    //   up   (blank q -> p): d'_p = 0, d'_q = d_p - 1 - less, d'_i = d_i - [a_i < t]
    //   down (blank p -> q): d'_p = d_q + 1 + less, d'_q = 0, d'_i = d_i + [a_i < t]
    // so the child's rank is the parent's plus the weighted digit changes
    // (w_i = (N-1-i)!, walked down from w_q by w_i = w_{i+1} * (N-1-i)).
    // an: the tiles of the permutation as nibbles, dn: the digits as nibbles.
    __device__ __forceinline__ static R vertical(R rank, uint64_t an, uint64_t dn, int blank, int width, bool up) {
        const int p = up ? blank - width : blank;
        const int q = up ? blank : blank + width;
        const uint32_t t = (uint32_t)((an >> (4 * (up ? p : q))) & 15u);
        R wq = 0;
#pragma unroll
        for (int i = 0; i < N; ++i) wq = (i == q) ? (R)kFact[N - 1 - i] : wq;
        R w = wq, sum = 0;
        uint32_t less = 0;
        for (int i = q - 1; i > p; --i) {
            w *= (R)(N - 1 - i);
            if ((uint32_t)((an >> (4 * i)) & 15u) < t) {
                ++less;
                sum += w;
            }
        }
        const R wp = w * (R)(N - 1 - p);
        if (up) {
            const R dp = (R)((dn >> (4 * p)) & 15u);
            return rank - dp * wp - sum + (dp - 1 - (R)less) * wq;
        }
        const R dq = (R)((dn >> (4 * q)) & 15u);
        return rank + ((R)less + 1 + dq) * wp + sum - dq * wq;
    }

    template <class F>
    __device__ __forceinline__ static void expand(uint64_t index, uint32_t op_bits, const PuzzleParams& prm,
                                                  F&& f) {
        const int width = prm.size;
        const R rank = (R)index;
        uint32_t d[N];
        digits(rank, d);
        // the blank (tile 0) is the leftmost position whose digit is 0
        int blank = N - 1;
#pragma unroll
        for (int p = N - 1; p >= 0; --p) blank = (d[p] == 0) ? p : blank;
        const int col = blank % width;
        const bool up = blank >= width && !(op_bits & (1u << OP_UP));
        const bool dn = blank < N - width && !(op_bits & (1u << OP_DOWN));
        const bool lf = col > 0 && !(op_bits & (1u << OP_LEFT));
        const bool rt = col < width - 1 && !(op_bits & (1u << OP_RIGHT));
        if (up || dn) {
            static_assert(N <= 16, "tiles and digits as nibbles of 64 bits");
            uint64_t tiles = 0xFEDCBA9876543210ull, an = 0, dn_ = 0;
#pragma unroll
            for (int i = 0; i < N; ++i) {  // digits -> tiles, as nibbles
                const uint32_t sh = d[i] * 4;
                an |= ((tiles >> sh) & 15ull) << (4 * i);
                dn_ |= (uint64_t)d[i] << (4 * i);
                tiles = ((tiles >> sh >> 4) << sh) | (tiles & ((1ull << sh) - 1ull));
            }
            if (up) f(0, (uint64_t)vertical(rank, an, dn_, blank, width, true), OP_DOWN);
            if (dn) f(1, (uint64_t)vertical(rank, an, dn_, blank, width, false), OP_UP);
        }
        // horizontal moves touch two adjacent digits; w_p = (N-1-p)!
        uint32_t dl = 0, dr = 0;
        R wl = 0, wb = 0, wr = 0;
#pragma unroll
        for (int p = 0; p < N; ++p) {
            dl = (p == blank - 1) ? d[p] : dl;
            dr = (p == blank + 1) ? d[p] : dr;
            wl = (p == blank - 1) ? (R)kFact[N - 1 - p] : wl;
            wb = (p == blank) ? (R)kFact[N - 1 - p] : wb;
            wr = (p == blank + 1) ? (R)kFact[N - 1 - p] : wr;
        }
        // blank at b moves left: digit at b-1 becomes 0, digit at b becomes d[b-1]-1
        if (lf) f(2, (uint64_t)(rank - (R)dl * wl + (R)(dl - 1) * wb), OP_RIGHT);
        // blank at b moves right: digit at b becomes d[b+1]+1, digit at b+1 becomes 0
        if (rt) f(3, (uint64_t)(rank + (R)(dr + 1) * wb - (R)dr * wr), OP_LEFT);
    }
};

// ---------------------------------------------------------- sliding (packed)

// SlidingTilePuzzleOptimized<W,H>: the index of a state is its tile list (the
// non-blank tiles in cell order) with each position's packed value -- the
// number of smaller tiles to its left, in [0, position] -- as mixed-radix
// digits of positions 2..11, the values of positions 12..14 as the segment's
// nibbles and the blank cell in the low four bits; positions 0 and 1 are
// implied by parity (the index format of Puzzles/SlidingPuzzleGpu.cu:142-159
// and Puzzles/SlidingTilePuzzleOptimized.cpp:79-192).
template <int W, int H>
struct Packed {
    static constexpr int OPS = 4;
    static constexpr int SLOTS = 4;
    static constexpr int S = W * H;

    // The index as 4-bit nibbles: nibble q = packed value of list position q
    // for q = 2..14 (positions 12..14 are the segment's nibbles), nibbles 0
    // and 1 zero, nibble 15 the blank cell.
    __device__ __forceinline__ static uint64_t nibbles(uint64_t index) {
        uint32_t idx = (uint32_t)index;
        uint64_t nib = ((uint64_t)(idx & 15u) << 60) | (((index >> 32) & 0xFFFull) << 48);
        idx >>= 4;
#pragma unroll
        for (int i = 2; i < 12; ++i) {
            const uint32_t q = idx / (uint32_t)(i + 1);
            nib |= (uint64_t)(idx - q * (uint32_t)(i + 1)) << (4 * i);
            idx = q;
        }
        return nib;
    }

    __device__ __forceinline__ static uint64_t index_of(uint64_t nib) {
        uint32_t idx = 0;
#pragma unroll
        for (int i = 11; i >= 2; --i) idx = idx * (uint32_t)(i + 1) + (uint32_t)((nib >> (4 * i)) & 15u);
        return (((nib >> 48) & 0xFFFull) << 32) | (uint64_t)(idx * 16u + (uint32_t)(nib >> 60));
    }

    // Children of the vertical moves without unpacking the permutation.  A
    // vertical move rotates W consecutive tile-list positions, so only their
    // packed values change, and by the relative order of those W tiles alone:
    // the unpack step applied inside the window to the window's packed values
    // gives the tiles' ranks within the list prefix ending at the window.
    // Position 0's packed value is 0 and position 1's is the parity bit that
    // unpack restores (the number of inversions is sum(i - v_i)).  Slot and
    // label order as expand: up (slot 0, OP_DOWN), then down (slot 1, OP_UP).
    template <class F>
    __device__ __forceinline__ static void vertical(uint64_t index, bool up, bool dn, F&& f) {
        uint64_t nib = nibbles(index);
        const int z = (int)(nib >> 60);
        constexpr uint32_t C = (uint32_t)((S - 2) * (S - 1) / 2 - 1);  // sum of i over 2..S-2
        const uint32_t inv = (uint32_t)__popcll(nib & 0x0111111111111100ull) + 1u + C;
        const bool inv_even = (inv & 1u) == 0;
        const bool row_even = ((z / W) & 1) == 0;
        const bool restore = (W % 2 == 0) ? ((!inv_even) == row_even) : !inv_even;
        nib |= (uint64_t)(restore ? 1u : 0u) << 4;
        constexpr uint64_t WMASK = (W == 16) ? ~0ull : ((1ull << (4 * W)) - 1ull);
        if (up) {  // the tile at position z-W moves to z-1, the W-1 after it move down one
            const int p = z - W;
            const uint64_t win = nib >> (4 * p);
            uint32_t v[W], R[W];
#pragma unroll
            for (int k = 0; k < W; ++k) R[k] = v[k] = (uint32_t)((win >> (4 * k)) & 15u);
#pragma unroll
            for (int i = 1; i < W; ++i)
#pragma unroll
                for (int j = 0; j < i; ++j) R[j] += (R[j] >= R[i]) ? 1u : 0u;
            uint64_t w = 0;
            uint32_t below = 0;
#pragma unroll
            for (int k = 1; k < W; ++k) {
                w |= (uint64_t)(v[k] - (R[0] < R[k] ? 1u : 0u)) << (4 * (k - 1));
                below += (R[k] < R[0]) ? 1u : 0u;
            }
            w |= (uint64_t)(v[0] + below) << (4 * (W - 1));
            uint64_t c = (nib & ~(WMASK << (4 * p))) | (w << (4 * p));
            c = (c & ~(15ull << 60)) | ((uint64_t)(z - W) << 60);
            f(0, index_of(c), OP_DOWN);
        }
        if (dn) {  // the tile at position z+W-1 moves to z, the W-1 before it move up one
            const int p = z;
            const uint64_t win = nib >> (4 * p);
            uint32_t v[W], R[W];
#pragma unroll
            for (int k = 0; k < W; ++k) R[k] = v[k] = (uint32_t)((win >> (4 * k)) & 15u);
#pragma unroll
            for (int i = 1; i < W; ++i)
#pragma unroll
                for (int j = 0; j < i; ++j) R[j] += (R[j] >= R[i]) ? 1u : 0u;
            uint64_t w = 0;
            uint32_t below = 0;
#pragma unroll
            for (int k = 0; k < W - 1; ++k) {
                w |= (uint64_t)(v[k] + (R[W - 1] < R[k] ? 1u : 0u)) << (4 * (k + 1));
                below += (R[k] < R[W - 1]) ? 1u : 0u;
            }
            w |= (uint64_t)(v[W - 1] - below);
            uint64_t c = (nib & ~(WMASK << (4 * p))) | (w << (4 * p));
            c = (c & ~(15ull << 60)) | ((uint64_t)(z + W) << 60);
            f(1, index_of(c), OP_UP);
        }
    }

    // Slots whose child may lie outside the parent's segment of 2^seg_bits
    // indexes.  A horizontal move changes only the blank cell, index -+ 1
    // without carry, so it stays in any segment of 16 or more indexes.  With
    // the 15-puzzle's own segments (seg_bits 32) the segment is the packed
    // values of tile-list positions 12..14, which name the last three tiles of
    // the list in order (the value of position 14 is its tile, that of 13 its
    // tile less one if tile 14 is smaller, ...).  A vertical move rotates W
    // consecutive list positions -- up from blank cell z positions z-W .. z-1,
    // down z .. z+W-1 -- and puts another tile at the highest of them.  So it
    // leaves the segment exactly when that position is 12 or more.
    __device__ __forceinline__ static uint32_t cross_slots(uint64_t index, int seg_bits) {
        const int z = (int)(index & 15u);
        uint32_t m = seg_bits >= 4 ? 0x3u : 0xFu;
        if (S == 16 && seg_bits == 32) {
            if (z - 1 < 12) m &= ~0x1u;
            if (z + W - 1 < 12) m &= ~0x2u;
        }
        return m;
    }

    // Slots whose child may lie inside the parent's segment: with the
    // 15-puzzle's own segments, the complement of cross_slots (exact, above);
    // otherwise every slot.
    __device__ __forceinline__ static uint32_t in_slots(uint64_t index, int seg_bits) {
        return (S == 16 && seg_bits == 32) ? (~cross_slots(index, seg_bits) & 0xFu) : 0xFu;
    }

    // slots that have a child: the move exists and its op bit is not set
    __device__ __forceinline__ static uint32_t live_slots(uint64_t index, uint32_t op_bits) {
        const int z = (int)(index & 15u);
        uint32_t m = 0;
        if (z >= W && !(op_bits & (1u << OP_UP))) m |= 1u;
        if (z < S - W && !(op_bits & (1u << OP_DOWN))) m |= 2u;
        if (z % W > 0 && !(op_bits & (1u << OP_LEFT))) m |= 4u;
        if (z % W < W - 1 && !(op_bits & (1u << OP_RIGHT))) m |= 8u;
        return m;
    }

    // `want`: the slots to compute (bit k: slot k); the others are skipped,
    // and the permutation is unpacked only when a vertical move is wanted.
    template <class F>
    __device__ __forceinline__ static void expand(uint64_t index, uint32_t op_bits, const PuzzleParams&, F&& f,
                                                  uint32_t want = 0xFu) {
        const int blank = (int)(index & 15u);
        const bool up = (want & 1u) && blank >= W && !(op_bits & (1u << OP_UP));
        const bool dn = (want & 2u) && blank < S - W && !(op_bits & (1u << OP_DOWN));
        if (up || dn) vertical(index, up, dn, f);
        if ((want & 4u) && blank % W > 0 && !(op_bits & (1u << OP_LEFT))) f(2, index - 1, OP_RIGHT);
        if ((want & 8u) && blank % W < W - 1 && !(op_bits & (1u << OP_RIGHT))) f(3, index + 1, OP_LEFT);
    }
};

// ------------------------------------------------------------ 4-peg Hanoi

// FourPegHanoiGPU / FourPegHanoiOptimizedGPU (Puzzles/HanoiTowersGPU.cu:114-144,
// 250-281): both kernels produce the same slots; one implementation serves both.
template <class U>
struct Hanoi {
    static constexpr int OPS = 4;
    static constexpr int SLOTS = 6;

    __device__ __forceinline__ static int ctz_(U x) {
        if constexpr (sizeof(U) == 4) return __ffs((int)x) - 1;
        else return __ffsll((unsigned long long)x) - 1;
    }
    __device__ __forceinline__ static int top_bit(U x) {
        if constexpr (sizeof(U) == 4) return x ? 31 - __clz((int)x) : -1;
        else return x ? 63 - __clzll((long long)x) : -1;
    }

    template <class F>
    __device__ __forceinline__ static void expand(uint64_t index64, uint32_t op_bits, const PuzzleParams& prm, F&& f) {
        const U index = (U)index64;
        const int disks = prm.size;
        const U m55 = (U)0x5555555555555555ull;
        const U bm = (2 * disks >= (int)(8 * sizeof(U))) ? m55 : (U)(((U)1 << (2 * disks)) - 1) & m55;
        const U p0 = index >> 1, p1 = index;
        U pg[4];
        pg[0] = ~p0 & ~p1 & bm;
        pg[1] = ~p0 & p1 & bm;
        pg[2] = p0 & ~p1 & bm;
        pg[3] = p0 & p1 & bm;
        int top[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) top[p] = pg[p] ? ctz_(pg[p]) : 255;
        const bool sym = (prm.flags & 1) != 0;
        int slot = 0;
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = a + 1; b < 4; ++b, ++slot) {
                const bool fwd = top[a] < top[b];
                const int src = fwd ? a : b, dst = fwd ? b : a;
                const int ts = fwd ? top[a] : top[b];
                const int td = fwd ? top[b] : top[a];
                if (op_bits & (1u << src)) continue;
                if (!(ts < td)) continue;  // both pegs empty
                U child = index ^ ((U)(src ^ dst) << ts);
                uint32_t label = (uint32_t)dst;
                if (sym) {
                    // sort pegs 1..3 of the child by bottom disk, largest first
                    U q[4];
#pragma unroll
                    for (int p = 0; p < 4; ++p) q[p] = pg[p];
                    const U bit = (U)1 << ts;
#pragma unroll
                    for (int p = 0; p < 4; ++p) {
                        q[p] = (p == src) ? (q[p] & ~bit) : q[p];
                        q[p] = (p == dst) ? (q[p] | bit) : q[p];
                    }
                    int bot[4], lab[4] = {0, 1, 2, 3};
#pragma unroll
                    for (int p = 0; p < 4; ++p) bot[p] = top_bit(q[p]);
                    const int order[3][2] = {{2, 3}, {1, 2}, {2, 3}};
#pragma unroll
                    for (int o = 0; o < 3; ++o) {
                        const int i = order[o][0], j = order[o][1];
                        if (bot[i] < bot[j]) {
                            int t = bot[i]; bot[i] = bot[j]; bot[j] = t;
                            U u = q[i]; q[i] = q[j]; q[j] = u;
                            t = lab[i]; lab[i] = lab[j]; lab[j] = t;
                        }
                    }
                    child = q[1] | (q[2] << 1) | (q[3] | (q[3] << 1));
                    uint32_t nl = 0;
#pragma unroll
                    for (int p = 0; p < 4; ++p) nl = (lab[p] == dst) ? (uint32_t)p : nl;
                    label = nl;
                }
                f(slot, (uint64_t)child, label);
            }
    }
};

}  // namespace psd
