// CPU check of the cross-tile list arithmetic of the 10..12-pancake search
// (puzzlesearch_amd/csrc/pancake_lists.h), the same functions the device
// kernels run, against the oracle's rank/unrank:
//   * every child of every parent of the checked tiles equals the oracle's
//     flip of the unranked parent, in slot order, with in-tile children
//     exactly the flips of <= 9 elements;
//   * every cross child's slot is in range, its list's destination tile is
//     the child's tile, and a whole tile fills each list exactly to capacity;
//   * the lists the apply pass enumerates for each destination tile are
//     exactly the lists that reach it, every list once.
// Usage: pancake_lists_check N tile [tile ...]   (exit status 0 = pass)
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "pancake_lists.h"

extern "C" {
uint64_t or_pancake_rank(int* arr, int n);
void or_pancake_unrank(uint64_t index, int* arr, int n);
}

namespace {

int g_fail = 0;

#define CHECK(cond, ...)                    \
    do {                                    \
        if (!(cond)) {                      \
            if (g_fail < 20) {              \
                std::printf("FAIL: ");      \
                std::printf(__VA_ARGS__);   \
                std::printf("\n");          \
            }                               \
            ++g_fail;                       \
        }                                   \
    } while (0)

// the table-driven routine (children_tab) visits exactly what children() visits
template <int N>
void check_tab(uint32_t T, uint32_t step) {
    std::vector<pkl::Tab6> tab6(720);
    std::vector<uint32_t> tab6lo(720), tab3(504);
    for (uint32_t l = 0; l < 720; ++l) tab6[l] = pkl::tab6_entry(l, &tab6lo[l]);
    for (uint32_t h = 0; h < 504; ++h) tab3[h] = pkl::tab3_entry(h);
    uint32_t dp[3];
    pkl::tile_digits<N>(T, dp);
    const pkl::TileK tk = pkl::tile_consts<N>(dp);
    for (uint32_t off = 0; off < pkl::kTile; off += step) {
        std::vector<uint64_t> a, b;
        pkl::children<N>(T, dp, off, [&](uint32_t c) { a.push_back(c); },
                         [&](int k, uint32_t slot, uint32_t c) { a.push_back(((uint64_t)k << 48) | ((uint64_t)slot << 32) | c); });
        pkl::children_tab<N>(tab6, tab6lo, tab3, tk, off, [&](uint32_t c) { b.push_back(c); },
                             [&](int k, uint32_t slot, uint32_t c) { b.push_back(((uint64_t)k << 48) | ((uint64_t)slot << 32) | c); });
        CHECK(a == b, "tile %u off %u: table-driven children differ (%zu vs %zu visits)", T, off, a.size(), b.size());
        if (a != b && g_fail < 5) {
            for (size_t i = 0; i < a.size() && i < b.size(); ++i)
                if (a[i] != b[i]) std::printf("  visit %zu: %llx vs %llx\n", i, (unsigned long long)a[i], (unsigned long long)b[i]);
        }
    }
}

template <int N>
void check_tile(uint32_t T) {
    using G = pkl::Geo<N>;
    uint32_t dp[3];
    pkl::tile_digits<N>(T, dp);
    check_tab<N>(T, 1);
    std::vector<uint32_t> fill(G::NSLOT, 0);
    for (uint32_t off = 0; off < pkl::kTile; ++off) {
        const uint64_t parent = (uint64_t)T * pkl::kTile + off;
        int a[16];
        or_pancake_unrank(parent, a, N);
        int k_expect = 2;  // slot order: flip 2, 3, .., N
        pkl::children<N>(
            T, dp, off,
            [&](uint32_t child_off) {
                int b[16];
                for (int i = 0; i < N; ++i) b[i] = a[i];
                for (int i = 0; i < k_expect / 2; ++i) {
                    const int t = b[N - k_expect + i];
                    b[N - k_expect + i] = b[N - 1 - i];
                    b[N - 1 - i] = t;
                }
                const uint64_t want = or_pancake_rank(b, N);
                CHECK(k_expect <= 9, "tile %u off %u: flip %d reported in-tile", T, off, k_expect);
                CHECK(want == (uint64_t)T * pkl::kTile + child_off, "tile %u off %u flip %d: child %llu, oracle %llu",
                      T, off, k_expect, (unsigned long long)((uint64_t)T * pkl::kTile + child_off),
                      (unsigned long long)want);
                ++k_expect;
            },
            [&](int k, uint32_t slot, uint32_t child) {
                int b[16];
                for (int i = 0; i < N; ++i) b[i] = a[i];
                for (int i = 0; i < k / 2; ++i) {
                    const int t = b[N - k + i];
                    b[N - k + i] = b[N - 1 - i];
                    b[N - 1 - i] = t;
                }
                const uint64_t want = or_pancake_rank(b, N);
                CHECK(k == k_expect && k >= 10, "tile %u off %u: cross flip %d out of order", T, off, k);
                CHECK(want == child, "tile %u off %u flip %d: child %u, oracle %llu", T, off, k, child,
                      (unsigned long long)want);
                CHECK(slot < pkl::kSlots[k - 10], "tile %u off %u flip %d: slot %u out of range", T, off, k, slot);
                if (slot < pkl::kSlots[k - 10]) {
                    CHECK(pkl::dest_of<N>(T, k, slot) == child / pkl::kTile,
                          "tile %u off %u flip %d slot %u: list goes to tile %u, child in tile %u", T, off, k, slot,
                          pkl::dest_of<N>(T, k, slot), child / pkl::kTile);
                    ++fill[pkl::kSlotBase[k - 10] + slot];
                }
                ++k_expect;
            });
        CHECK(k_expect == N + 1, "tile %u off %u: %d children", T, off, k_expect - 2);
    }
    for (int k = 10; k <= N; ++k)
        for (uint32_t s = 0; s < pkl::kSlots[k - 10]; ++s)
            CHECK(fill[pkl::kSlotBase[k - 10] + s] == pkl::kCap[k - 10], "tile %u flip %d slot %u: %u entries, cap %u",
                  T, k, s, fill[pkl::kSlotBase[k - 10] + s], pkl::kCap[k - 10]);
}

template <int N>
void check_incoming() {
    using G = pkl::Geo<N>;
    std::vector<uint8_t> seen((size_t)G::TILES * G::NSLOT, 0);
    for (uint32_t t = 0; t < G::TILES; ++t) {
        uint32_t x[3];
        pkl::tile_prefix<N>(t, x);
        CHECK(pkl::tile_of_prefix<N>(x) == t, "tile %u: prefix round trip", t);
        for (uint32_t i = 0; i < G::NSLOT; ++i) {
            int k = 0;
            uint32_t src = 0, slot = 0;
            pkl::incoming<N>(t, i, &k, &src, &slot);
            CHECK(src < G::TILES && slot < pkl::kSlots[k - 10], "tile %u list %u: source %u slot %u", t, i, src, slot);
            if (src >= G::TILES || slot >= pkl::kSlots[k - 10]) continue;
            CHECK(pkl::dest_of<N>(src, k, slot) == t, "tile %u list %u: (%u, %d, %u) goes to %u", t, i, src, k, slot,
                  pkl::dest_of<N>(src, k, slot));
            const size_t id = (size_t)src * G::NSLOT + pkl::kSlotBase[k - 10] + slot;
            CHECK(!seen[id], "list (%u, %d, %u) enumerated twice", src, k, slot);
            seen[id] = 1;
            CHECK(pkl::list_base(src, G::CROSS, k, slot) + pkl::kCap[k - 10] <= G::LIST_ENTRIES,
                  "list (%u, %d, %u) ends beyond the list array", src, k, slot);
        }
    }
    size_t n = 0;
    for (auto v : seen) n += v;
    CHECK(n == seen.size(), "%zu of %zu lists enumerated", n, seen.size());
}

template <int N>
int run(int argc, char** argv) {
    check_incoming<N>();
    for (int i = 2; i < argc; ++i) {
        const uint32_t T = (uint32_t)std::strtoul(argv[i], nullptr, 10);
        if (T < pkl::Geo<N>::TILES) check_tile<N>(T);
    }
    std::printf("pancake %d: %d failures\n", N, g_fail);
    return g_fail ? 1 : 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::printf("usage: %s N tile [tile ...]\n", argv[0]);
        return 2;
    }
    switch (std::atoi(argv[1])) {
    case 10: return run<10>(argc, argv);
    case 11: return run<11>(argc, argv);
    case 12: return run<12>(argc, argv);
    }
    std::printf("N must be 10, 11 or 12\n");
    return 2;
}
