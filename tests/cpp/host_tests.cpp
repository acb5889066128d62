// Tests of the C++ host mirror (puzzlesearch_amd/host) written like the
// reference's GoogleTest cases (PuzzlesTests/*.cpp): the same puzzles, calls
// and expected values, with a minimal EXPECT framework (GoogleTest is not in
// this image).
//
//   host_tests --cpu   tests that need no GPU (Parse/ToString/metadata)
//   host_tests --gpu   everything (plugin Expand on the GPU, the drivers)
#include <cstring>
#include <functional>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "InMemoryBFS.h"
#include "Puzzles.h"

namespace {

int g_failures = 0;
int g_checks = 0;
std::string g_current;

#define EXPECT_EQ(a, b)                                                                                     \
    do {                                                                                                    \
        ++g_checks;                                                                                         \
        auto _va = (a);                                                                                     \
        auto _vb = (b);                                                                                     \
        if (!(_va == _vb)) {                                                                                \
            ++g_failures;                                                                                   \
            std::cerr << "FAILED " << g_current << " " << __FILE__ << ":" << __LINE__ << ": " #a " == " #b   \
                      << "\n   got: " << _va << "\n  want: " << _vb << std::endl;                            \
        }                                                                                                   \
    } while (0)
#define EXPECT_NE(a, b)                                                                                     \
    do {                                                                                                    \
        ++g_checks;                                                                                         \
        if ((a) == (b)) {                                                                                   \
            ++g_failures;                                                                                   \
            std::cerr << "FAILED " << g_current << " " << __FILE__ << ":" << __LINE__ << ": " #a " != " #b   \
                      << std::endl;                                                                         \
        }                                                                                                   \
    } while (0)

struct Case {
    const char* name;
    bool gpu;
    std::function<void()> fn;
};
std::vector<Case>& cases() {
    static std::vector<Case> c;
    return c;
}
struct Reg {
    Reg(const char* n, bool g, std::function<void()> f) { cases().push_back({n, g, f}); }
};
#define TEST_CPU(name) \
    static void name(); \
    static Reg reg_##name(#name, false, name); \
    static void name()
#define TEST_GPU(name) \
    static void name(); \
    static Reg reg_##name(#name, true, name); \
    static void name()

std::string Join(const std::vector<uint64_t>& v) {
    std::ostringstream s;
    for (size_t i = 0; i < v.size(); i++) s << (i ? " " : "") << v[i];
    return s.str();
}

struct Expanded {
    std::vector<uint64_t> children;
    std::vector<int> ops;
};

Expanded ExpandOne(Puzzle& puz, uint64_t index, ExpandHint hint = {}) {
    std::vector<uint64_t> indexes{index};
    std::vector<int> used{0};
    Expanded e;
    puz.Expand(indexes, used, e.children, e.ops, hint);
    return e;
}

// golden vectors (layer counts), with their reference source
const char* ST3x2 = "1 2 3 5 6 7 10 12 12 16 23 25 28 39 44 40 29 21 18 12 6 1";  // PuzzlesTests/PuzzleSearchTests.cpp:20
const char* ST4x2 =                                                                   // PuzzlesTests/PuzzleSearchTests.cpp:21
    "1 2 3 6 10 14 19 28 42 61 85 119 161 215 293 396 506 632 788 985 1194 1414 1664 1884 1999 1958 1770 1463 1076 "
    "667 361 190 88 39 19 7 1";
const char* PANCAKE5 = "1 4 12 35 48 20";                                               // PuzzleSearchTests.cpp:544
const char* PANCAKE9 = "1 8 56 391 2278 10666 38015 93585 132697 79379 5804";           // PuzzleSearchTests.cpp:550
const char* HANOI10_SYM =                                                              // PuzzleSearchTests.cpp:372
    "1 1 1 3 5 7 11 16 24 36 55 56 81 137 172 162 183 293 435 506 588 562 779 1056 1519 1780 1930 1983 2249 2871 "
    "3957 5367 6165 6638 6425 6654 7608 9511 12500 15255 17690 18320 15339 10630 7614 5496 2214 334 52 4";

// ------------------------------------------------------------- CPU only

TEST_CPU(PancakeParse) {
    PancakeOptimizedGPU p5(5, false);
    EXPECT_EQ(p5.ToString(0), std::string("0 1 2 3 4"));
    EXPECT_EQ(p5.Parse("0 1 2 3 4"), 0ull);
    PancakeOptimizedGPU p7(7, false);
    EXPECT_EQ(p7.ToString(p7.Parse("5 1 3 6 0 2 4")), std::string("5 1 3 6 0 2 4"));
    PancakeOptimizedGPU p15(15, true);
    EXPECT_EQ(p15.Parse("0 1 2 3 4 5 6 7 8 9 10 11 12 13 14"), 0ull);
    EXPECT_EQ(p15.ToString(0), std::string("0 1 2 3 4 5 6 7 8 9 10 11 12 13 14"));
    EXPECT_EQ(p15.ToString(p15.Parse("14 13 12 11 10 9 8 7 6 5 4 3 2 1 0")),
              std::string("14 13 12 11 10 9 8 7 6 5 4 3 2 1 0"));
    EXPECT_EQ(PancakeOptimizedGPU(12, false).IndexesCount(), 1ull << 29);  // PancakeOptimized.cpp:94-105
    EXPECT_EQ(PancakeOptimizedGPU(11, false).IndexesCount(), 39916800ull);
    EXPECT_EQ(PancakeOptimizedGPU(13, false).IndexesCount(), 13ull << 29);
}

TEST_CPU(PermutationRankKat) {
    // PuzzlesTests/PermutationMapTest.cpp:81-94
    int a[] = {8, 9, 10, 11, 0, 1, 2, 3, 4, 5, 6, 7};
    EXPECT_EQ(ps_host::PermutationRank(a, 12), 351590400ull);
    int b[] = {5, 1, 2, 4, 0, 3};
    EXPECT_EQ(ps_host::PermutationRank(b, 6), 634ull);
    int c[16];
    ps_host::PermutationUnrank(20922789888000ull - 1, c, 16);
    EXPECT_EQ(c[0], 15);
    EXPECT_EQ(c[15], 0);
}

TEST_CPU(SlidingParse) {
    SlidingTilePuzzleGpu s(3, 2);
    EXPECT_EQ(s.Parse("0 1 2 3 4 5"), 0ull);
    EXPECT_EQ(s.ToString(s.Parse("3 1 2 0 4 5")), std::string("3 1 2 0 4 5"));
    EXPECT_EQ(s.IndexesCount(), 720ull);
    SlidingTilePuzzleOptimized<4, 4> o;
    const std::string st = "0 1 2 3 4 5 6 7 8 9 10 11 12 13 14 15";
    EXPECT_EQ(o.ToString(o.Parse(st)), st);
    EXPECT_EQ(o.IndexesCount(), 3805ull << 32);
    SlidingTilePuzzleOptimized<4, 3> o43;
    EXPECT_EQ(o43.IndexesCount(), 319334400ull);
    const std::string s43 = "4 1 2 3 0 5 6 7 8 9 10 11";  // blank moved down once
    EXPECT_EQ(o43.ToString(o43.Parse(s43)), s43);
}

TEST_CPU(HanoiParse) {
    // PuzzlesTests/FourPegHanoiTest.cpp:15-18, 36-39
    FourPegHanoiGPU h5(5, false);
    EXPECT_EQ(h5.Parse("1 2 3 4 5 0 0 0"), 0ull);
    EXPECT_EQ(h5.ToString(0), std::string("1 2 3 4 5 0 0 0"));
    FourPegHanoiOptimizedGPU h7(7, false);
    EXPECT_EQ(h7.Parse("1 5 6 0 2 0 0 3 4 7"), 12532ull);
    EXPECT_EQ(h7.ToString(12532), std::string("1 5 6 0 2 0 0 3 4 7"));
    EXPECT_EQ(FourPegHanoiGPU(16, false).IndexesCount(), 1ull << 32);
}

// ------------------------------------------------------------------ GPU

TEST_GPU(OptPancake5GPU) {
    // PuzzlesTests/PancakeTest.cpp:176-197
    PancakeOptimizedGPU puz(5, false);
    auto e = ExpandOne(puz, 0);
    EXPECT_EQ(puz.ToString(e.children[0]), std::string("0 1 2 4 3"));
    EXPECT_EQ(e.ops[0], 0);
    EXPECT_EQ(puz.ToString(e.children[1]), std::string("0 1 4 3 2"));
    EXPECT_EQ(e.ops[1], 1);
    EXPECT_EQ(puz.ToString(e.children[2]), std::string("0 4 3 2 1"));
    EXPECT_EQ(e.ops[2], 2);
    EXPECT_EQ(puz.ToString(e.children[3]), std::string("4 3 2 1 0"));
    EXPECT_EQ(e.ops[3], 3);
}

TEST_GPU(OptPancake7GPU) {
    // PuzzlesTests/PancakeTest.cpp:199-223
    PancakeOptimizedGPU puz(7, false);
    auto e = ExpandOne(puz, puz.Parse("5 1 3 6 0 2 4"));
    const char* want[] = {"5 1 3 6 0 4 2", "5 1 3 6 4 2 0", "5 1 3 4 2 0 6",
                          "5 1 4 2 0 6 3", "5 4 2 0 6 3 1", "4 2 0 6 3 1 5"};
    for (int i = 0; i < 6; i++) {
        EXPECT_EQ(puz.ToString(e.children[i]), std::string(want[i]));
        EXPECT_EQ(e.ops[i], i);
    }
}

TEST_GPU(OptPancakeGpu15) {
    // PuzzlesTests/PancakeTest.cpp:225-283 (and :285-344 with the inverse index)
    for (bool inv : {false, true}) {
        PancakeOptimizedGPU puz(15, inv);
        auto e = ExpandOne(puz, 0);
        EXPECT_EQ(e.children.size(), 14u);
        EXPECT_EQ(puz.ToString(e.children[0]), std::string("0 1 2 3 4 5 6 7 8 9 10 11 12 14 13"));
        EXPECT_EQ(puz.ToString(e.children[13]), std::string("14 13 12 11 10 9 8 7 6 5 4 3 2 1 0"));
        for (int i = 0; i < 14; i++) EXPECT_EQ(e.ops[i], i);
        for (int i = 0; i < 11; i++) EXPECT_EQ(e.children[i] >> 29, 0ull);
        EXPECT_EQ(e.children[11] >> 29, 12ull);
        EXPECT_EQ(e.children[12] >> 29, 181ull);
        EXPECT_EQ(e.children[13] >> 29, 2729ull);
    }
}

TEST_GPU(Pancake15SegmentHints) {
    // PancakeOptimizedGPU::Expand with 29-bit segment hints: in-segment ops 0..10,
    // cross-segment ops 11..13 (Puzzles/PancakeOptimizedGpu.cpp:57-79)
    PancakeOptimizedGPU puz(15, false);
    auto in = ExpandOne(puz, 0, ExpandHint{29, false});
    auto xs = ExpandOne(puz, 0, ExpandHint{29, true});
    EXPECT_EQ(in.children.size(), 11u);
    EXPECT_EQ(xs.children.size(), 3u);
    EXPECT_EQ(xs.ops[0], 11);
    EXPECT_EQ(xs.children[2] >> 29, 2729ull);
}

TEST_GPU(SlidingGpu) {
    // PuzzlesTests/SlidingPuzzleTest.cpp:43-78
    SlidingTilePuzzleGpu puz(3, 2);
    auto e = ExpandOne(puz, puz.Parse("0 1 2 3 4 5"));
    EXPECT_EQ(e.children[0], Puzzle::INVALID_INDEX);
    EXPECT_EQ(e.ops[0], 15);
    EXPECT_EQ(puz.ToString(e.children[1]), std::string("3 1 2 0 4 5"));
    EXPECT_NE(e.ops[1], 15);
    EXPECT_EQ(e.children[2], Puzzle::INVALID_INDEX);
    EXPECT_EQ(puz.ToString(e.children[3]), std::string("1 0 2 3 4 5"));
    SlidingTilePuzzleGpu p44(4, 4);
    auto f = ExpandOne(p44, p44.Parse("0 1 2 3 4 5 6 7 8 9 10 11 12 13 14 15"));
    EXPECT_EQ(p44.ToString(f.children[1]), std::string("4 1 2 3 0 5 6 7 8 9 10 11 12 13 14 15"));
    EXPECT_EQ(p44.ToString(f.children[3]), std::string("1 0 2 3 4 5 6 7 8 9 10 11 12 13 14 15"));
}

TEST_GPU(SlidingOptimized) {
    // PuzzlesTests/SlidingPuzzleTest.cpp:80-115
    SlidingTilePuzzleOptimized<3, 2> puz;
    auto e = ExpandOne(puz, puz.Parse("0 1 2 3 4 5"));
    EXPECT_EQ(e.children[0], Puzzle::INVALID_INDEX);
    EXPECT_EQ(e.ops[0], 15);
    EXPECT_EQ(puz.ToString(e.children[1]), std::string("3 1 2 0 4 5"));
    EXPECT_EQ(puz.ToString(e.children[3]), std::string("1 0 2 3 4 5"));
    SlidingTilePuzzleOptimized<4, 4> p44;
    auto f = ExpandOne(p44, p44.Parse("0 1 2 3 4 5 6 7 8 9 10 11 12 13 14 15"));
    EXPECT_EQ(p44.ToString(f.children[1]), std::string("4 1 2 3 0 5 6 7 8 9 10 11 12 13 14 15"));
    EXPECT_EQ(p44.ToString(f.children[3]), std::string("1 0 2 3 4 5 6 7 8 9 10 11 12 13 14 15"));
}

TEST_GPU(HanoiGpu) {
    // PuzzlesTests/FourPegHanoiTest.cpp:245-337
    FourPegHanoiOptimizedGPU h7(7, false);
    auto e = ExpandOne(h7, h7.Parse("1 5 6 0 2 0 0 3 4 7"));
    const char* want[] = {"5 6 0 1 2 0 0 3 4 7", "5 6 0 2 0 1 0 3 4 7", "5 6 0 2 0 0 1 3 4 7",
                          "1 5 6 0 0 2 0 3 4 7", "1 5 6 0 0 0 2 3 4 7", "1 5 6 0 2 0 3 0 4 7"};
    const int wops[] = {1, 2, 3, 2, 3, 2};
    for (int i = 0; i < 6; i++) {
        EXPECT_EQ(h7.ToString(e.children[i]), std::string(want[i]));
        EXPECT_EQ(e.ops[i], wops[i]);
    }
    FourPegHanoiGPU s7(7, true);
    auto f = ExpandOne(s7, s7.Parse("3 4 6 0 1 5 7 0 2 0"));
    const char* wsym[] = {"1 3 4 6 0 5 7 0 2 0", "2 3 4 6 0 1 5 7 0 0", "4 6 0 1 5 7 0 3 0 2",
                          "3 4 6 0 5 7 0 1 2 0", "3 4 6 0 5 7 0 2 0 1", "3 4 6 0 1 5 7 0 2 0"};
    const int wsops[] = {0, 0, 2, 2, 3, 2};
    for (int i = 0; i < 6; i++) {
        EXPECT_EQ(s7.ToString(f.children[i]), std::string(wsym[i]));
        EXPECT_EQ(f.ops[i], wsops[i]);
    }
}

// PuzzlesTests/PuzzleSearchTests.cpp:97-119 (GPU plugins in the in-memory drivers)
TEST_GPU(SlidingTile_GPU_TwoBitBFS) {
    SlidingTilePuzzleGpu p32(3, 2);
    EXPECT_EQ(Join(InMemoryTwoBitBFS(p32, "0 1 2 3 4 5")), std::string(ST3x2));
    SlidingTilePuzzleGpu p42(4, 2);
    EXPECT_EQ(Join(InMemoryTwoBitBFS(p42, "0 1 2 3 4 5 6 7")), std::string(ST4x2));
}

TEST_GPU(SlidingTile_GPU_ThreeBitBFS) {
    SlidingTilePuzzleGpu p32(3, 2);
    EXPECT_EQ(Join(InMemoryThreeBitBFS(p32, "0 1 2 3 4 5")), std::string(ST3x2));
    SlidingTilePuzzleGpu p42(4, 2);
    EXPECT_EQ(Join(InMemoryThreeBitBFS(p42, "0 1 2 3 4 5 6 7")), std::string(ST4x2));
}

TEST_GPU(Pancake_GPU_ThreeBitBFS) {
    PancakeOptimizedGPU p5(5, false), p9(9, false);
    EXPECT_EQ(Join(InMemoryThreeBitBFS(p5, p5.ToString(0))), std::string(PANCAKE5));
    EXPECT_EQ(Join(InMemoryThreeBitBFS(p9, p9.ToString(0))), std::string(PANCAKE9));
}

TEST_GPU(Hanoi_GPU_ThreeBitBFS) {
    FourPegHanoiGPU h(10, true);
    EXPECT_EQ(Join(InMemoryThreeBitBFS(h, h.ToString(0))), std::string(HANOI10_SYM));
}

TEST_GPU(DeviceResidentDrivers) {
    PancakeOptimizedGPU p9(9, false);
    for (int seg : {0, 12, 15}) {
        for (bool batched : {true, false}) {
            GpuSearchOptions o;
            o.segmentBits = seg;
            o.batched = batched;
            EXPECT_EQ(Join(GpuInMemoryThreeBitBFS(p9, p9.ToString(0), o)), std::string(PANCAKE9));
        }
    }
    EXPECT_EQ(Join(GpuInMemoryTwoBitBFS(p9, p9.ToString(0))), std::string(PANCAKE9));
    SlidingTilePuzzleGpu p42(4, 2);
    EXPECT_EQ(Join(GpuInMemoryThreeBitBFS(p42, "0 1 2 3 4 5 6 7")), std::string(ST4x2));
    FourPegHanoiOptimizedGPU h(10, true);
    EXPECT_EQ(Join(GpuInMemoryThreeBitBFS(h, h.ToString(0))), std::string(HANOI10_SYM));
    GpuSearchOptions o;
    o.maxSteps = 3;
    EXPECT_EQ(Join(GpuInMemoryThreeBitBFS(p9, p9.ToString(0), o)), std::string("1 8 56 391"));
}

TEST_GPU(ErrorsBecomeExceptions) {
    // a start state outside the index space surfaces as std::runtime_error
    PancakeOptimizedGPU p6(6, false);
    GpuSearchOptions o;
    o.indexCount = 10;
    bool thrown = false;
    try {
        GpuInMemoryThreeBitBFS(p6, "5 4 3 2 1 0", o);
    } catch (const std::runtime_error&) {
        thrown = true;
    }
    EXPECT_EQ(thrown, true);
}

}  // namespace

int main(int argc, char** argv) {
    const bool gpu = argc > 1 && std::strcmp(argv[1], "--gpu") == 0;
    int run = 0;
    for (auto& c : cases()) {
        if (c.gpu && !gpu) continue;
        g_current = c.name;
        const int before = g_failures;
        try {
            c.fn();
        } catch (const std::exception& e) {
            ++g_failures;
            std::cerr << "FAILED " << c.name << ": exception " << e.what() << std::endl;
        }
        std::cout << (g_failures == before ? "[  OK  ] " : "[ FAIL ] ") << c.name << std::endl;
        ++run;
    }
    std::cout << run << " tests, " << g_checks << " checks, " << g_failures << " failures" << std::endl;
    return g_failures ? 1 : 0;
}
