// CPU check of the packed sliding-tile move generator (psd::Packed<W,H> of
// puzzlesearch_amd/csrc/device_puzzles.h, compiled for the host): every slot
// of random states of every shape, with random used-operator bits, equals the
// oracle's slot (or_expand_slots, the reference's GPU slot layout:
// Puzzles/SlidingPuzzleGpu.cu:210-270 with the unpack/rotate/pack steps of
// :108-207).  Any combination of packed digits and blank cell is a state (the
// parity of positions 0 and 1 is implied), so states are drawn digit by digit.
// The simple Lehmer-rank encoding (psd::Sliding<N>), whose vertical moves
// update the rank from the digits of the swapped window, is checked the same
// way on nine board shapes.
// Usage: packed_check count seed   (exit status 0 = pass)
#include "simt_emu.h"

#include <cinttypes>
#include <cstdio>
#include <random>
#include <vector>

#include "device_puzzles.h"
#include "oracle.h"

namespace {

int g_fail = 0;

template <int W, int H>
void check(uint64_t count, uint64_t seed) {
    using P = psd::Packed<W, H>;
    constexpr int S = W * H;
    std::mt19937_64 rng(seed * 1000003ull + (uint64_t)(W * 16 + H));
    std::vector<uint64_t> in(count), want(count * 4), got(count * 4, ~0ull);
    for (uint64_t i = 0; i < count; ++i) {
        uint64_t idx = 0;
        for (int q = 11; q >= 2; --q) idx = idx * (uint64_t)(q + 1) + (q <= S - 2 ? rng() % (uint64_t)(q + 1) : 0);
        uint64_t seg = 0;
        for (int q = 14; q >= 12; --q) seg = (seg << 4) | (q <= S - 2 ? rng() % (uint64_t)(q + 1) : 0);
        const uint64_t z = rng() % (uint64_t)S;
        const uint64_t index = (seg << 32) | (idx * 16 + z);
        in[i] = (index << 4) | (rng() % 16);
    }
    const or_puzzle op{OR_SLIDING_PACKED, W, H, 0};
    if (or_expand_slots(&op, OR_EXPAND_FULL, in.data(), count, want.data()) != 4) {
        std::printf("FAIL: oracle has no packed %dx%d\n", W, H);
        ++g_fail;
        return;
    }
    const psd::PuzzleParams prm{OR_SLIDING_PACKED, W, H, 0};
    for (uint64_t i = 0; i < count; ++i)
        P::expand(in[i] >> 4, (uint32_t)(in[i] & 15u), prm,
                  [&](int slot, uint64_t child, uint32_t label) { got[i * 4 + slot] = (child << 4) | label; });
    uint64_t bad = 0;
    for (uint64_t i = 0; i < count * 4; ++i)
        if (got[i] != want[i]) {
            if (bad < 5)
                std::printf("  %dx%d parent %" PRIx64 " slot %d: %" PRIx64 " vs oracle %" PRIx64 "\n", W, H, in[i / 4] >> 4,
                            (int)(i % 4), got[i], want[i]);
            ++bad;
        }
    std::printf("packed %dx%d: %" PRIu64 " parents, %" PRIu64 " slots differ\n", W, H, count, bad);
    g_fail += bad != 0;
}

// psd::Sliding<N> (the simple Lehmer-rank encoding, SlidingTilePuzzleSimple
// and the GPU kernel of Puzzles/SlidingPuzzleGpu.cu): random ranks below N!
// with random used-operator bits against the oracle's slots
template <int W, int H>
void check_simple(uint64_t count, uint64_t seed) {
    using P = psd::Sliding<W * H>;
    constexpr int N = W * H;
    uint64_t nfact = 1;
    for (int i = 2; i <= N; ++i) nfact *= (uint64_t)i;
    std::mt19937_64 rng(seed * 7000003ull + (uint64_t)(W * 16 + H));
    std::vector<uint64_t> in(count), want(count * 4), got(count * 4, ~0ull);
    for (uint64_t i = 0; i < count; ++i) in[i] = ((rng() % nfact) << 4) | (rng() % 16);
    const or_puzzle op{OR_SLIDING, W, H, 0};
    if (or_expand_slots(&op, OR_EXPAND_FULL, in.data(), count, want.data()) != 4) {
        std::printf("FAIL: oracle has no simple %dx%d\n", W, H);
        ++g_fail;
        return;
    }
    const psd::PuzzleParams prm{OR_SLIDING, W, H, 0};
    for (uint64_t i = 0; i < count; ++i)
        P::expand(in[i] >> 4, (uint32_t)(in[i] & 15u), prm,
                  [&](int slot, uint64_t child, uint32_t label) { got[i * 4 + slot] = (child << 4) | label; });
    uint64_t bad = 0;
    for (uint64_t i = 0; i < count * 4; ++i)
        if (got[i] != want[i]) {
            if (bad < 5)
                std::printf("  simple %dx%d parent %" PRIx64 " slot %d: %" PRIx64 " vs oracle %" PRIx64 "\n", W, H,
                            in[i / 4] >> 4, (int)(i % 4), got[i], want[i]);
            ++bad;
        }
    std::printf("simple %dx%d: %" PRIu64 " parents, %" PRIu64 " slots differ\n", W, H, count, bad);
    g_fail += bad != 0;
}

}  // namespace

int main(int argc, char** argv) {
    const uint64_t count = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 100000;
    const uint64_t seed = argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 1;
    check<2, 2>(count, seed);
    check<3, 2>(count, seed);
    check<4, 2>(count, seed);
    check<5, 2>(count, seed);
    check<6, 2>(count, seed);
    check<7, 2>(count, seed);
    check<8, 2>(count, seed);
    check<3, 3>(count, seed);
    check<4, 3>(count, seed);
    check<5, 3>(count, seed);
    check<4, 4>(count, seed);
    check_simple<2, 2>(count, seed);
    check_simple<3, 2>(count, seed);
    check_simple<4, 2>(count, seed);
    check_simple<3, 3>(count, seed);
    check_simple<5, 2>(count, seed);
    check_simple<4, 3>(count, seed);
    check_simple<6, 2>(count, seed);
    check_simple<7, 2>(count, seed);
    check_simple<4, 4>(count, seed);
    std::printf("%d shapes failed\n", g_fail);
    return g_fail ? 1 : 0;
}
