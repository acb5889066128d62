// Cross-tile child lists of the three-bit search of 10, 11 and 12 pancakes:
// index arithmetic shared by the device kernels and the CPU checks.
//
// Ranks are the lexicographic ranks of Puzzles/PancakeOptimized.cpp:54-77 for
// n <= 12 (digit i has radix N-i).  The index space is cut into tiles of 9!
// consecutive ranks: a tile fixes the first P = N-9 elements (its prefix) and
// holds every order of the other 9.  Flip k (reversal of the last k elements,
// op k-2 of Puzzles/PancakeGpu.cu:92-118) changes only the last k digits, so
// flips of k <= 9 keep the child in the parent's tile.  A flip of k >= 10
// moves the parent's last k-9 elements into the prefix: the child keeps the
// parent's first s = N-k elements and continues with the parent's elements at
// positions N-1, N-2, ...  Its tile is therefore fixed by the relative ranks,
// among the parent's 9 non-prefix elements, of the parent's last k-9
// elements: 9!/(18-k)! possible tiles ("slots"), each reached by exactly
// (18-k)! parents of the tile.  The children of flip k of source tile T thus
// go to 9!/(18-k)! lists of capacity (18-k)!, 9! entries per (T, k) in all,
// filled only by the workgroup that expands T.
//
// The apply pass takes a destination tile t and enumerates the lists that
// can reach it: for flip k the source tiles share t's first s prefix
// elements and hold any (k-9)-arrangement of t's 9 non-prefix elements in
// prefix positions s..P-1; the slot is given by the relative ranks of t's
// prefix elements s..P-1 within the source's non-prefix elements.
#pragma once

#include <cstdint>

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define PKL_HD __host__ __device__
#define PKL_UNROLL _Pragma("unroll")
#else
#define PKL_HD
#define PKL_UNROLL
#endif

namespace pkl {

constexpr uint32_t kTile = 362880u;       // 9! ranks per tile
constexpr uint32_t kTileWords = 5670u;    // 9! / 64 bit-array words per tile
constexpr uint32_t kFact[13] = {1u,      1u,       2u,        6u,         24u,        120u,      720u,
                                5040u,   40320u,   362880u,   3628800u,   39916800u,  479001600u};
// per cross flip k = 10, 11, 12 (index k - 10)
constexpr uint32_t kSlots[3] = {9u, 72u, 504u};      // destination tiles of one source tile
constexpr uint32_t kCap[3] = {40320u, 5040u, 720u};  // parents of one source tile per destination tile
constexpr uint32_t kSlotBase[4] = {0u, 9u, 81u, 585u};

template <int N>
struct Geo {
    static_assert(N >= 10 && N <= 12, "cross-tile lists cover 10..12 pancakes");
    static constexpr int P = N - 9;      // prefix elements fixed in a tile
    static constexpr int CROSS = N - 9;  // cross flips k = 10..N
    static constexpr uint32_t NSLOT = kSlotBase[N - 9];
    static constexpr uint32_t TILES = kFact[N] / kTile;  // 10, 110, 1320
    static constexpr uint64_t LIST_ENTRIES = (uint64_t)TILES * CROSS * kTile;
};

// c * f for a count c < 2^8 and a factorial constant f (24-bit multiply on the
// device when f fits in 24 bits)
PKL_HD inline uint32_t mul_small(uint32_t c, uint32_t f) {
#if defined(__HIP_DEVICE_COMPILE__)
    return (f < (1u << 24)) ? __umul24(c, f) : c * f;
#else
    return c * f;
#endif
}

// Slot of flip k from the relative ranks r1, r2, r3 (within the parent's
// non-prefix elements) of the parent's elements at positions N-1, N-2, N-3:
// mixed radix 9, 8, 7 over the ranks with earlier choices removed.
PKL_HD inline uint32_t slot_of(int k, uint32_t r1, uint32_t r2, uint32_t r3) {
    const uint32_t a2 = r2 - (r2 > r1 ? 1u : 0u);
    const uint32_t a3 = r3 - (r3 > r1 ? 1u : 0u) - (r3 > r2 ? 1u : 0u);
    return k == 10 ? r1 : (k == 11 ? r1 * 8u + a2 : (r1 * 8u + a2) * 7u + a3);
}

// Lehmer digits of a tile's prefix (digit i has radix N-i).
template <int N>
PKL_HD inline void tile_digits(uint32_t t, uint32_t (&dp)[3]) {
    constexpr int P = N - 9;
    // tile id = sum_{i < P} d_i * (N-1-i)!/9!: mixed radix N, N-1, .., N-P+1
    uint32_t r = t;
    PKL_UNROLL
    for (int i = 2; i >= 0; --i) {
        if (i < P) {
            const uint32_t m = (uint32_t)(N - i);
            dp[i] = r % m;
            r /= m;
        } else {
            dp[i] = 0;
        }
    }
}

// Visits the children of one parent of tile t (rank t*9! + off, off < 9!) in
// slot order: in(off') for flips of k <= 9, where off' is the child's offset
// in the same tile, and cross(k, slot, child) for flips of k >= 10.  dp are
// the tile's prefix digits (tile_digits).  Arithmetic: the right-to-left sweep
// over the Lehmer digits of the pancake functor (device_puzzles.h) in 32 bits.
template <int N, class FI, class FX>
PKL_HD inline void children(uint32_t t, const uint32_t (&dp)[3], uint32_t off, FI&& in, FX&& cross) {
    constexpr int P = N - 9;
    uint32_t d[N];
    PKL_UNROLL
    for (int i = 0; i < P; ++i) d[i] = dp[i];
    uint32_t r = off;
    PKL_UNROLL
    for (int i = N - 1; i >= P; --i) {
        const uint32_t m = (uint32_t)(N - i);
        d[i] = r % m;
        r /= m;
    }
    const uint32_t index = t * kTile + off;
    // cc[q] = #{p in [s, q) : a[p] < a[q]} for the current suffix start s;
    // a[s] < a[q]  <=>  cc[q] (for s+1) + d[q] >= d[s]
    uint32_t cc[N];
    PKL_UNROLL
    for (int i = 0; i < N; ++i) cc[i] = 0;
    uint32_t suffix = 0;  // sum_{i >= s} d[i] * (N-1-i)!
    uint32_t r1 = 0, r2 = 0, r3 = 0;
    PKL_UNROLL
    for (int s = N - 2; s >= 0; --s) {
        const int k = N - s;
        uint32_t rho = 0;  // rank of the reversed suffix pattern
        PKL_UNROLL
        for (int q = s + 1; q < N; ++q) {
            cc[q] += (cc[q] + d[q] >= d[s]) ? 1u : 0u;
            rho += mul_small(cc[q], kFact[q - s]);
        }
        suffix += mul_small(d[s], kFact[N - 1 - s]);
        if (s == P) {
            // relative ranks of a[N-1], a[N-2], a[N-3] among the non-prefix elements
            r1 = cc[N - 1] + d[N - 1];
            r2 = cc[N - 2] + d[N - 2];
            r3 = cc[N - 3] + d[N - 3];
        }
        if (k <= 9) in(off - suffix + rho);
        else cross(k, slot_of(k, r1, r2, r3), index - suffix + rho);
    }
}

// k-th smallest element of the set given as a bit mask (k < popcount)
PKL_HD inline uint32_t select_elem(uint32_t mask, uint32_t k) {
    uint32_t e = 0;
    PKL_UNROLL
    for (int b = 0; b < 12; ++b) {
        if ((mask >> b) & 1u) {
            if (k == 0) {
                e = (uint32_t)b;
                k = ~0u;  // found; stop counting
            } else if (k != ~0u) {
                --k;
            }
        }
    }
    return e;
}

// rank of element e among the elements of mask (number of mask elements < e)
PKL_HD inline uint32_t rank_in(uint32_t mask, uint32_t e) {
    const uint32_t below = mask & ((1u << e) - 1u);
#if defined(__HIP_DEVICE_COMPILE__)
    return (uint32_t)__popc(below);
#else
    return (uint32_t)__builtin_popcount(below);
#endif
}

// ---------------------------------------------------------------------------
// Table-driven form of children() (same visits, same order).
//
// Positions are numbered from the end here: slot j is position N-1-j, and the
// per-position values (digits d, running counts cc) are kept four to a 32-bit
// word, one byte per slot (word w holds slots 4w..4w+3).
//
//  * The last six positions: a 720-entry table indexed by l = off mod 6!
//    gives their digits, the tile offsets of the flips of 2..6 elements
//    (relative to off - l) and the counts cc after the sweep reaches position
//    N-6 (the flips of <= 6 elements involve only these six elements).
//  * The next three positions: a 504-entry table indexed by off / 6! gives
//    their digits.
//  * The sweep over positions s = N-7..0 updates all active slots at once:
//    cc[q] += (cc[q] + d[q] >= d[s]) is a byte-wise compare done by adding
//    0x80 - d[s] to each byte of cc + d and taking bit 7.  The rank of the
//    reversed suffix, sum_q cc[q] (q-s)!, is a byte dot product (v_dot4) of
//    the count word with factorial weights; weights above 255 are factored
//    as (common factorial) x (ratio <= 255).

struct alignas(16) Tab6 {
    uint32_t d0;   // digits of slots 0..3
    uint32_t cc0;  // counts of slots 0..3 after position N-6
    uint32_t c234; // offsets (< 720) of the flips of 2, 3, 4 elements, 10 bits each
    uint32_t c56;  // flips of 5 and 6 elements (10 bits each), bit 20: count of slot 4
};

// Entry l of the six-position table (pattern of the last six elements with
// Lehmer digits e0..e5, radices 6..1); *lo receives the digits of slots 4, 5.
PKL_HD inline Tab6 tab6_entry(uint32_t l, uint32_t* lo) {
    uint32_t e[6], r = l;
    for (int i = 5; i >= 0; --i) {
        e[i] = r % (uint32_t)(6 - i);
        r /= (uint32_t)(6 - i);
    }
    uint32_t a[6], avail = 0x3Fu;
    for (int i = 0; i < 6; ++i) {
        a[i] = select_elem(avail, e[i]);
        avail &= ~(1u << a[i]);
    }
    auto rank6 = [](const uint32_t(&b)[6]) {
        uint32_t seen = 0, rr = 0;
        for (int i = 0; i < 6; ++i) {
            rr = rr * (uint32_t)(6 - i) + (b[i] - rank_in(seen, b[i]) - 0u);
            seen |= 1u << b[i];
        }
        return rr;
    };
    uint32_t c[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int k = 2; k <= 6; ++k) {
        uint32_t b[6];
        for (int i = 0; i < 6; ++i) b[i] = a[i];
        for (int i = 0; i < k / 2; ++i) {
            const uint32_t t = b[6 - k + i];
            b[6 - k + i] = b[5 - i];
            b[5 - i] = t;
        }
        c[k] = rank6(b);
    }
    // counts c(0, q) = #{p in [0, q) : a[p] < a[q]} of the six-pattern; slot j is pattern index 5-j
    uint32_t cq[6];
    for (int q = 0; q < 6; ++q) {
        cq[q] = 0;
        for (int p2 = 0; p2 < q; ++p2) cq[q] += a[p2] < a[q] ? 1u : 0u;
    }
    Tab6 t;
    t.d0 = e[5] | (e[4] << 8) | (e[3] << 16) | (e[2] << 24);
    t.cc0 = cq[5] | (cq[4] << 8) | (cq[3] << 16) | (cq[2] << 24);
    t.c234 = c[2] | (c[3] << 10) | (c[4] << 20);
    t.c56 = c[5] | (c[6] << 10) | (cq[1] << 20);
    *lo = e[1] | (e[0] << 8);
    return t;
}

// Entry h of the three-position table: the digits of positions N-7, N-8, N-9
// (radices 7, 8, 9) of h = off / 6!, one per byte.
PKL_HD inline uint32_t tab3_entry(uint32_t h) {
    const uint32_t d7 = h % 7u, t = h / 7u;
    return d7 | ((t % 8u) << 8) | ((t / 8u) << 16);
}

// Per-tile constants of children_tab().
struct TileK {
    uint32_t td2;     // digits of slots 9..11 (the tile prefix) in bytes 1..3
    uint32_t pre[3];  // sum_{i < s} d_i (N-1-i)! for s = 0, 1, 2
};

template <int N>
PKL_HD inline TileK tile_consts(const uint32_t (&dp)[3]) {
    constexpr int P = N - 9;
    TileK k;
    k.td2 = 0;
    for (int i = 0; i < P; ++i) k.td2 |= dp[i] << (8 * (N - 1 - i - 8));  // position i is slot N-1-i
    k.pre[0] = 0;
    k.pre[1] = dp[0] * kFact[N - 1];
    k.pre[2] = k.pre[1] + (P > 1 ? dp[1] * kFact[N - 2] : 0u);
    return k;
}

PKL_HD inline uint32_t udot4(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_udot4(a, b, c, false);
#else
    for (int i = 0; i < 4; ++i) c += ((a >> (8 * i)) & 255u) * ((b >> (8 * i)) & 255u);
    return c;
#endif
}

// byte b of w in all four bytes
PKL_HD inline uint32_t splat_byte(uint32_t w, int b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_perm(w, w, 0x01010101u * (uint32_t)b);
#else
    return ((w >> (8 * b)) & 255u) * 0x01010101u;
#endif
}

// Weight groups of one count word at sweep step m (= N-1-s active slots):
// slot j has weight (m-j)! (0 when j >= m); a group is a byte-weight word
// (ratios <= 255) and the common factor of its members.
struct WGroups {
    uint32_t pack[4];
    uint32_t factor[4];
    int n;
};

constexpr uint32_t cfact(int n) { return n <= 1 ? 1u : (uint32_t)n * cfact(n - 1); }

constexpr WGroups weight_groups(int m, int w) {
    WGroups g{{0, 0, 0, 0}, {0, 0, 0, 0}, 0};
    int b = 0;
    while (b < 4 && 4 * w + b < m) {
        // extend the group from byte b while the ratio to its smallest weight stays <= 255
        int e = b;
        while (e + 1 < 4 && 4 * w + e + 1 < m &&
               cfact(m - 4 * w - b) / cfact(m - 4 * w - (e + 1)) <= 255u)
            ++e;
        const uint32_t f = cfact(m - 4 * w - e);
        uint32_t pack = 0;
        for (int x = b; x <= e; ++x) pack |= (cfact(m - 4 * w - x) / f) << (8 * x);
        g.pack[g.n] = pack;
        g.factor[g.n] = f;
        ++g.n;
        b = e + 1;
    }
    return g;
}

// sum over the active slots of word w of cc[slot] * weight, added to acc
template <int M, int W>
PKL_HD inline uint32_t rho_word(uint32_t ccw, uint32_t acc) {
    constexpr WGroups g = weight_groups(M, W);
    PKL_UNROLL
    for (int i = 0; i < g.n; ++i) {
        if (g.factor[i] == 1u) acc = udot4(ccw, g.pack[i], acc);
        else acc += mul_small(udot4(ccw, g.pack[i], 0u), g.factor[i]);
    }
    return acc;
}

// mask of the low bit of every active byte of word w (slots < m)
constexpr uint32_t active_lsb(int m, int w) {
    uint32_t r = 0;
    for (int b = 0; b < 4; ++b)
        if (4 * w + b < m) r |= 1u << (8 * b);
    return r;
}

template <int N, int S>
struct Sweep {
    // one step of the sweep at position S (S = N-7 .. 0) and the steps below it
    template <class FI, class FX>
    PKL_HD static inline void run(uint32_t (&cc)[3], const uint32_t (&d)[3], const uint32_t (&e)[3],
                                  const uint32_t (&base)[3], const TileK& tk, uint32_t& r1, uint32_t& r2,
                                  uint32_t& r3, FI& in, FX& cross) {
        constexpr int P = N - 9;
        constexpr int M = N - 1 - S;  // active slots 0..M-1; position S is slot M
        const uint32_t ns = 0u - splat_byte(d[M >> 2], M & 3);
        PKL_UNROLL
        for (int w = 0; w < 3; ++w) {
            if (4 * w < M) {
                const uint32_t y = cc[w] + e[w] + ns;  // byte b: cc + d + 0x80 - d[S], no borrow
                cc[w] += (y >> 7) & active_lsb(M, w);
            }
        }
        uint32_t rho = rho_word<M, 0>(cc[0], 0u);
        rho = rho_word<M, 1>(cc[1], rho);
        if constexpr (M > 8) rho = rho_word<M, 2>(cc[2], rho);
        if constexpr (S == P) {
            const uint32_t x = cc[0] + d[0];
            r1 = x & 255u;
            r2 = (x >> 8) & 255u;
            r3 = (x >> 16) & 255u;
        }
        constexpr int K = N - S;
        if constexpr (K <= 9) in(base[S - P] + rho);
        else cross(K, slot_of(K, r1, r2, r3), tk.pre[S] + rho);
        if constexpr (S > 0) Sweep<N, S - 1>::run(cc, d, e, base, tk, r1, r2, r3, in, cross);
    }
};

// The table entries of one parent: tab6[l], tab6lo[l] and tab3[h] for the
// tile offset off = h*720 + l.
struct TabRow {
    Tab6 t;
    uint32_t lo, hd;
};
template <class TA, class TB, class TC>
PKL_HD inline TabRow tab_row(const TA& tab6, const TB& tab6lo, const TC& tab3, uint32_t off) {
    const uint32_t h = off / 720u;
    const uint32_t l = off - h * 720u;
    return TabRow{tab6[l], tab6lo[l], tab3[h]};
}

// children() from the parent's table row (tab_row of the same offset).
template <int N, class FI, class FX>
PKL_HD inline void children_row(const TabRow& row, const TileK& tk, uint32_t off, FI&& in, FX&& cross) {
    constexpr int P = N - 9;
    const uint32_t h = off / 720u;
    const uint32_t l = off - h * 720u;
    const Tab6& t = row.t;
    const uint32_t lo = row.lo;
    const uint32_t hd = row.hd;
    uint32_t d[3], cc[3], e[3];
    d[0] = t.d0;
    d[1] = lo | (hd << 16);             // slots 4, 5 | slots 6, 7 (positions N-7, N-8)
    d[2] = ((hd >> 16) & 255u) | tk.td2;  // slot 8 (position N-9) | the tile prefix
    cc[0] = t.cc0;
    cc[1] = (t.c56 >> 20) & 1u;
    cc[2] = 0;
    PKL_UNROLL
    for (int w = 0; w < 3; ++w) e[w] = d[w] + 0x80808080u;
    // flips of 2..6 elements: from the table
    const uint32_t b6 = off - l;
    in(b6 + (t.c234 & 1023u));
    in(b6 + ((t.c234 >> 10) & 1023u));
    in(b6 + ((t.c234 >> 20) & 1023u));
    in(b6 + (t.c56 & 1023u));
    in(b6 + ((t.c56 >> 10) & 1023u));
    // in-tile base of the flips of 9, 8, 7 elements: the tile offset less the
    // reversed suffix (index - suffix), by position s = P, P+1, P+2
    uint32_t base[3];
    base[0] = 0;
    base[1] = ((hd >> 16) & 255u) * 40320u;
    base[2] = (h - (hd & 255u)) * 720u;
    uint32_t r1 = 0, r2 = 0, r3 = 0;
    Sweep<N, N - 7>::run(cc, d, e, base, tk, r1, r2, r3, in, cross);
    (void)P;
}

// children() with the tables: tab6[l] (and tab6lo[l] = digits of slots 4, 5)
// for l < 720, tab3[h] for h < 504.
template <int N, class TA, class TB, class TC, class FI, class FX>
PKL_HD inline void children_tab(const TA& tab6, const TB& tab6lo, const TC& tab3, const TileK& tk, uint32_t off,
                                FI&& in, FX&& cross) {
    children_row<N>(tab_row(tab6, tab6lo, tab3, off), tk, off, in, cross);
}

// Prefix elements of tile t.
template <int N>
PKL_HD inline void tile_prefix(uint32_t t, uint32_t (&x)[3]) {
    constexpr int P = N - 9;
    uint32_t dp[3];
    tile_digits<N>(t, dp);
    uint32_t avail = (1u << N) - 1u;
    PKL_UNROLL
    for (int i = 0; i < 3; ++i) {
        if (i < P) {
            x[i] = select_elem(avail, dp[i]);
            avail &= ~(1u << x[i]);
        } else {
            x[i] = 0;
        }
    }
}

// Tile of a prefix given as elements.
template <int N>
PKL_HD inline uint32_t tile_of_prefix(const uint32_t (&x)[3]) {
    constexpr int P = N - 9;
    uint32_t avail = (1u << N) - 1u, t = 0;
    PKL_UNROLL
    for (int i = 0; i < 3; ++i) {
        if (i < P) {
            t = t * (uint32_t)(N - i) + rank_in(avail, x[i]);
            avail &= ~(1u << x[i]);
        }
    }
    return t;
}

// List i (0 <= i < NSLOT) arriving at destination tile t: flip k, source tile
// and slot.
template <int N>
PKL_HD inline void incoming(uint32_t t, uint32_t i, int* k_out, uint32_t* src, uint32_t* slot) {
    constexpr int P = N - 9;
    const int k = i < kSlotBase[1] ? 10 : (i < kSlotBase[2] ? 11 : 12);
    uint32_t q = i - kSlotBase[k - 10];
    const int s = N - k;   // prefix positions 0..s-1 are shared with t
    const int m = k - 9;   // prefix positions s..P-1 of the source: an m-arrangement of t's free elements
    uint32_t x[3];
    tile_prefix<N>(t, x);
    uint32_t free_el = (1u << N) - 1u;
    PKL_UNROLL
    for (int j = 0; j < 3; ++j)
        if (j < P) free_el &= ~(1u << x[j]);
    // q = mixed radix (9, 8, 7) over the first m choices
    uint32_t jsel[3] = {0, 0, 0};
    if (m == 1) {
        jsel[0] = q;
    } else if (m == 2) {
        jsel[0] = q / 8u;
        jsel[1] = q % 8u;
    } else {
        jsel[2] = q % 7u;
        q /= 7u;
        jsel[1] = q % 8u;
        jsel[0] = q / 8u;
    }
    uint32_t y[3];
    uint32_t pre = 0;  // the source's prefix elements
    PKL_UNROLL
    for (int j = 0; j < 3; ++j) {
        if (j < s) {
            y[j] = x[j];
            pre |= 1u << y[j];
        }
    }
    uint32_t avail = free_el;
    PKL_UNROLL
    for (int j = 0; j < 3; ++j) {
        if (j < m) {
            const uint32_t e = select_elem(avail, jsel[j]);
            avail &= ~(1u << e);
            y[s + j] = e;
            pre |= 1u << e;
        }
    }
    *src = tile_of_prefix<N>(y);
    // the source's parents at positions N-1, N-2, N-3 hold t's prefix elements s, s+1, s+2
    const uint32_t nonpre = ((1u << N) - 1u) & ~pre;
    const uint32_t r1 = rank_in(nonpre, x[s]);
    const uint32_t r2 = (m >= 2) ? rank_in(nonpre, x[s + 1 < 3 ? s + 1 : 2]) : 0u;
    const uint32_t r3 = (m >= 3) ? rank_in(nonpre, x[s + 2 < 3 ? s + 2 : 2]) : 0u;
    *slot = slot_of(k, r1, r2, r3);
    *k_out = k;
}

// Destination tile of list (source tile T, flip k, slot).
template <int N>
PKL_HD inline uint32_t dest_of(uint32_t T, int k, uint32_t slot) {
    constexpr int P = N - 9;
    const int s = N - k, m = k - 9;
    uint32_t y[3];
    tile_prefix<N>(T, y);
    uint32_t nonpre = (1u << N) - 1u;
    PKL_UNROLL
    for (int j = 0; j < 3; ++j)
        if (j < P) nonpre &= ~(1u << y[j]);
    uint32_t a[3] = {0, 0, 0};  // adjusted ranks
    if (m == 1) {
        a[0] = slot;
    } else if (m == 2) {
        a[0] = slot / 8u;
        a[1] = slot % 8u;
    } else {
        a[2] = slot % 7u;
        a[0] = (slot / 7u) / 8u;
        a[1] = (slot / 7u) % 8u;
    }
    uint32_t x[3];
    PKL_UNROLL
    for (int j = 0; j < 3; ++j)
        if (j < s) x[j] = y[j];
    uint32_t avail = nonpre;
    PKL_UNROLL
    for (int j = 0; j < 3; ++j) {
        if (j < m) {
            const uint32_t e = select_elem(avail, a[j]);
            avail &= ~(1u << e);
            x[s + j] = e;
        }
    }
    return tile_of_prefix<N>(x);
}

// First entry of list (T, k, slot) in the list array.
PKL_HD inline uint64_t list_base(uint32_t T, int cross, int k, uint32_t slot) {
    return ((uint64_t)T * (uint64_t)cross + (uint64_t)(k - 10)) * kTile + (uint64_t)slot * kCap[k - 10];
}

}  // namespace pkl
