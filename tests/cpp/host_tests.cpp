// Tests of the C++ host mirror (puzzlesearch_amd/host) written like the
// reference's GoogleTest cases (PuzzlesTests/*.cpp): the same puzzles, calls
// and expected values, with a minimal EXPECT framework (GoogleTest is not in
// this image).
//
//   host_tests --cpu   tests that need no GPU (Parse/ToString/metadata)
//   host_tests --gpu   everything (plugin Expand on the GPU, the drivers)
#include <algorithm>
#include <cstring>
#include <functional>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include <cstdlib>
#include <stdexcept>
#include <unistd.h>

#include "CompressedFrontier.h"
#include "DiskBasedBFS.h"
#include "HanoiSearch.h"
#include "HanoiTowers.h"
#include "InMemoryBFS.h"
#include "Puzzles.h"
#include "Store.h"

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
const char* ST5x2 =                                                                    // PuzzleSearchTests.cpp:22
    "1 2 3 6 11 19 30 44 68 112 176 271 411 602 851 1232 1783 2530 3567 4996 6838 9279 12463 16597 21848 28227 "
    "35682 44464 54597 65966 78433 91725 104896 116966 126335 131998 133107 128720 119332 106335 91545 75742 60119 "
    "45840 33422 23223 15140 9094 5073 2605 1224 528 225 75 20 2";
const char* PANCAKE13_6 = "1 12 132 1451 14556 130096 1030505";                          // PuzzleSearchTests.cpp:569
const char* HANOI21_10 = "1 3 6 12 30 30 66 96 126 210 330";                             // PuzzleSearchTests.cpp:477
const char* HANOI21_SYM_10 = "1 1 1 3 5 7 11 16 24 36 55";                               // PuzzleSearchTests.cpp:487
const char* HANOI17_120 =                                                               // HanoiSearchTests.cpp:21
    "1 3 6 12 30 30 66 96 126 210 330 318 462 816 1032 936 1044 1752 2610 3036 3528 3306 4578 6318 9108 "
    "10674 11580 11844 13374 17124 23664 32184 36984 39822 38544 39936 45936 57990 77262 96366 116052 "
    "127092 122142 112086 116814 135078 157326 192498 246468 304650 354018 388128 415206 427710 411804 "
    "402504 451236 521094 600684 714750 883602 1064970 1228656 1349322 1464336 1539426 1565634 1554408 "
    "1610640 1703436 1890282 2144742 2533656 2988834 3506376 4017900 4551204 4999416 5378250 5623770 "
    "5778096 5794866 5857608 6013998 6447654 7084050 7998492 9102018 10433148 11919366 13440804 14761050 "
    "16074066 17531574 18802386 19692462 19862262 19671618 19386408 19539486 20568270 22353276 24582186 "
    "27002898 29769948 32947008 36487218 40325658 44952132 49926018 54298296 57553218 59394084 60022506 "
    "59865378 60725346 63215580 65921856 68036658 69281274 70563468";

// a fresh directory for the stores of one test
std::string TempDir(const char* tag) {
    std::string t = std::string("/tmp/ps_host_tests_") + tag + "_XXXXXX";
    std::vector<char> buf(t.begin(), t.end());
    buf.push_back('\0');
    const char* d = mkdtemp(buf.data());
    if (!d) throw std::runtime_error("mkdtemp failed");
    return d;
}

void RemoveDir(const std::string& dir) { std::system(("rm -rf '" + dir + "'").c_str()); }

// CommonTests/StoreTests.cpp:9-37: 300 chunks of 10 KiB round-robin over the
// segments, read back three times after rewinds, nothing left after DeleteAll
void CheckStore(Store& store) {
    constexpr size_t SIZE = 10 * 1024;
    constexpr int BUFS = 300;
    const int segments = store.MaxSegments();
    std::vector<uint8_t> buf(SIZE);
    for (int i = 0; i < BUFS; i++) {
        std::memset(buf.data(), i & 255, SIZE);
        store.Write(i % segments, buf.data(), SIZE);
    }
    EXPECT_EQ(store.TotalLength(), (uint64_t)SIZE * BUFS);
    for (int tries = 0; tries < 3; tries++) {
        bool same = true;
        for (int i = 0; i < BUFS; i++) {
            const size_t read = store.Read(i % segments, buf.data(), SIZE);
            EXPECT_EQ(read, SIZE);
            for (size_t j = 0; j < SIZE; j++) same &= buf[j] == (uint8_t)(i & 255);
        }
        EXPECT_EQ(same, true);
        store.RewindAll();
    }
    store.DeleteAll();
    for (int i = 0; i < segments; i++) EXPECT_EQ(store.Read(i, buf.data(), SIZE), (size_t)0);
    EXPECT_EQ(store.TotalLength(), 0ull);
}

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

TEST_CPU(StoreLayouts) {
    // CommonTests/StoreTests.cpp:39-80: every file layout of StoreOptions
    const std::string dir = TempDir("store");
    {
        auto single = Store::CreateSingleFileStore(50, dir + "/file");
        CheckStore(single);
        auto multi = Store::CreateMultiFileStore(50, dir + "/dir", "store");
        CheckStore(multi);
        auto seq = Store::CreateSequentialStore(50, {dir + "/s1", dir + "/s2", dir + "/s3"});
        CheckStore(seq);
        for (int files : {0, 1, 5}) {
            StoreOptions opts;
            opts.directories = {dir + "/d1", dir + "/d2", dir + "/d3"};
            opts.filesPerPath = files;
            auto store = Store::CreateFileStore(50, "store" + std::to_string(files), opts);
            CheckStore(store);
        }
        // chunks keep their boundaries; Delete drops one segment only
        auto s = Store::CreateSingleFileStore(3, dir + "/chunks");
        const uint8_t a[3] = {1, 2, 3}, b[5] = {4, 5, 6, 7, 8};
        s.Write(1, a, 3);
        s.Write(2, b, 5);
        s.Write(1, b, 5);
        uint8_t r[16];
        EXPECT_EQ(s.Read(1, r, 16), (size_t)3);
        EXPECT_EQ((int)r[2], 3);
        EXPECT_EQ(s.Read(1, r, 16), (size_t)5);
        EXPECT_EQ((int)r[4], 8);
        EXPECT_EQ(s.Read(1, r, 16), (size_t)0);
        s.Delete(1);
        EXPECT_EQ(s.HasData(1), false);
        EXPECT_EQ(s.Length(2), 5ull);
        EXPECT_EQ(s.Read(2, r, 16), (size_t)5);
        StoreSet set;
        set.Stores.emplace_back(Store::CreateSingleFileStore(2, dir + "/x0"));
        set.Stores.emplace_back(Store::CreateSingleFileStore(2, dir + "/x1"));
        set.Stores[1].Write(0, a, 3);
        StoreSet other;
        other.Stores.emplace_back(Store::CreateSingleFileStore(2, dir + "/y0"));
        other.Stores.emplace_back(Store::CreateSingleFileStore(2, dir + "/y1"));
        set.Swap(other);
        EXPECT_EQ(other.TotalLength(), 3ull);
        EXPECT_EQ(set.TotalLength(), 0ull);
    }
    RemoveDir(dir);
}

TEST_CPU(FrontierStreamChunks) {
    // chunk lengths of FrontierCompression streams (Common/FrontierCompression.cpp:117-125):
    // a bitmap chunk of 2 words holding first=100 and 3 more bits, then a
    // StreamVInt chunk of 5 values (8 after padding: two tuples)
    std::vector<uint8_t> s;
    auto put32 = [&](uint32_t v) {
        for (int i = 0; i < 4; i++) s.push_back((uint8_t)(v >> (8 * i)));
    };
    put32(2u | 0x80000000u);
    put32(100);
    const uint64_t w0 = 0x5ull, w1 = 1ull << 63;
    for (uint64_t w : {w0, w1})
        for (int i = 0; i < 8; i++) s.push_back((uint8_t)(w >> (8 * i)));
    put32(5);
    s.push_back(0x1B);  // lengths 4, 3, 2, 1 bytes
    for (int i = 0; i < 10; i++) s.push_back(0);
    s.push_back(0x00);  // four 1-byte values
    for (int i = 0; i < 4; i++) s.push_back(0);
    std::vector<uint64_t> off;
    EXPECT_EQ(fcstream::ChunkBoundaries(s.data(), s.size(), off), 4ull + 5ull);
    EXPECT_EQ(off.size(), (size_t)3);
    EXPECT_EQ(off[1], 24ull);
    EXPECT_EQ(off[2], (uint64_t)s.size());
    bool thrown = false;
    try {
        fcstream::ChunkBoundaries(s.data(), s.size() - 1, off);
    } catch (const std::runtime_error&) {
        thrown = true;
    }
    EXPECT_EQ(thrown, true);
}

// ------------------------------------------------------------------ GPU

TEST_GPU(DiskBasedOptThreeBitBFS_Reference) {
    // PuzzlesTests/PuzzleSearchTests.cpp:258-333, 461-490, 553-590: the disk
    // driver over file Stores, with the segment steps on the device
    const std::string dir = TempDir("disk");
    auto run = [&](Puzzle& p, const std::string& init, int segBits, size_t maxSteps, int filesPerPath) {
        PuzzleOptions o;
        o.segmentBits = segBits;
        o.maxSteps = maxSteps;
        o.storeOptions.directories = {dir + "/a", dir + "/b"};
        o.storeOptions.filesPerPath = filesPerPath;
        return Join(DiskBasedOptThreeBitBFS(p, init, o));
    };
    SlidingTilePuzzleGpu s52(5, 2);
    EXPECT_EQ(run(s52, "0 1 2 3 4 5 6 7 8 9", 22, 10000, 5), std::string(ST5x2));
    EXPECT_EQ(run(s52, "0 1 2 3 4 5 6 7 8 9", 19, 10000, 0), std::string(ST5x2));
    SlidingTilePuzzleOptimized<5, 2> o52;
    EXPECT_EQ(run(o52, "0 1 2 3 4 5 6 7 8 9", 22, 10000, 1), std::string(ST5x2));
    EXPECT_EQ(run(o52, "0 1 2 3 4 5 6 7 8 9", 19, 10000, 5), std::string(ST5x2));
    PancakeOptimizedGPU p9(9, false);
    EXPECT_EQ(run(p9, p9.ToString(0), 15, 10000, 5), std::string(PANCAKE9));
    PancakeOptimizedGPU p5(5, false);
    EXPECT_EQ(run(p5, p5.ToString(0), 32, 10000, 5), std::string(PANCAKE5));
    FourPegHanoiGPU h10(10, true);
    EXPECT_EQ(run(h10, h10.ToString(0), 20, 10000, 5), std::string(HANOI10_SYM));
    FourPegHanoiOptimizedGPU ho10(10, true);
    EXPECT_EQ(run(ho10, ho10.ToString(0), 14, 10000, 0), std::string(HANOI10_SYM));
    for (bool inv : {false, true}) {
        PancakeOptimizedGPU p13(13, inv);
        EXPECT_EQ(run(p13, p13.ToString(0), 29, 6, 5), std::string(PANCAKE13_6));
    }
    FourPegHanoiGPU h21(21, false), h21s(21, true);
    EXPECT_EQ(run(h21, h21.ToString(0), 32, 10, 5), std::string(HANOI21_10));
    EXPECT_EQ(run(h21s, h21s.ToString(0), 32, 10, 5), std::string(HANOI21_SYM_10));
    RemoveDir(dir);
}

TEST_GPU(CompressedSegmentReaderWriter) {
    // CompressedSegmentWriter's chunks read back by CompressedSegmentReader
    // chunk by chunk (Common/CompressedFrontier.cpp:90-158), across Store pieces
    const std::string dir = TempDir("frontier");
    {
        auto store = Store::CreateSingleFileStore(4, dir + "/f");
        CompressedSegmentWriter w(store);
        w.SetSegment(2);
        std::vector<uint32_t> ids;
        for (uint32_t i = 0; i < 700000; i++) ids.push_back(i * 3 + (i % 7 == 0));  // ascending
        for (uint32_t v : ids) w.Add(v);
        w.Flush();
        CompressedSegmentReader r(store);
        r.SetSegment(2);
        std::vector<uint32_t> back;
        for (;;) {
            auto& b = r.Read();
            if (b.IsEmpty()) break;
            EXPECT_EQ(b.Size() <= (size_t)(1u << 18), true);
            back.insert(back.end(), b.Buf(), b.Buf() + b.Size());
        }
        EXPECT_EQ(back == ids, true);
        r.SetSegment(1);
        EXPECT_EQ(r.Read().IsEmpty(), true);
    }
    RemoveDir(dir);
}

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

TEST_CPU(HanoiTowersVerticalMeta) {
    // HanoiTowers/HanoiTowers.h:13-23, HanoiTowers.cpp:139-150
    using H17 = HanoiTowers<17>;
    EXPECT_EQ(H17::IndexesCount(), 1ull << 34);
    EXPECT_EQ(H17::MaxSegments(), 4);
    EXPECT_EQ(HanoiTowers<24>::MaxSegments(), 65536);
    EXPECT_EQ(H17::ToString(0), std::string("1 2 3 4 5 6 7 8 9 10 11 12 13 14 15 16 17 0 0 0"));
    EXPECT_EQ(H17::Parse(H17::ToString(0)), 0ull);
    const uint64_t x = (1ull << 32) | (2ull << 2) | 1ull;  // disk 17 on peg 1, disk 2 on peg 2, disk 1 on peg 1
    EXPECT_EQ(H17::ToString(x), std::string("3 4 5 6 7 8 9 10 11 12 13 14 15 16 0 1 17 0 2 0"));
    EXPECT_EQ(H17::Parse(H17::ToString(x)), x);
    EXPECT_EQ(H17::SplitIndex(x).first, 1);
    EXPECT_EQ(H17::SplitIndex(x).second, (uint32_t)9);
    EXPECT_EQ(H17::PegsMult(x), 6);
    EXPECT_EQ(H17::PegsMult(1ull << 32), 3);  // only peg 1 used besides peg 0
    EXPECT_EQ(H17::InSegmentOnly(0), false);
    EXPECT_EQ(H17::InSegmentOnly(0x1B), true);  // disks 1..3 on pegs 3, 2, 1
}

TEST_GPU(HanoiTowersVerticalExpand) {
    // slot form (HanoiExpandGPU.cu:98-229) against the full expansion
    using H = HanoiTowers<17>;
    std::vector<uint64_t> frontier{0}, all{0};
    for (int layer = 0; layer < 12; layer++) {
        std::vector<uint64_t> next;
        H::Expand(frontier, next);
        frontier.clear();
        for (uint64_t c : next)
            if (std::find(all.begin(), all.end(), c) == all.end()) {
                all.push_back(c);
                frontier.push_back(c);
            }
    }
    std::vector<uint32_t> ids;
    for (uint64_t s : all) ids.push_back(uint32_t(s));
    std::vector<uint32_t> six, three;
    H::ExpandInSegment(0, ids.size(), ids.data(), six);
    GpuExpandInSegmentWithoutSmallest<17>(0, ids.size(), ids.data(), three);
    EXPECT_EQ(six.size(), ids.size() * 6);
    EXPECT_EQ(three.size(), ids.size() * 3);
    std::vector<uint64_t> full;
    H::Expand(all, full);
    bool same = true;
    for (size_t i = 0; i < ids.size(); i++)
        for (int k = 0; k < 6; k++) {
            const uint64_t c = full[i * 6 + k];
            same &= six[i * 6 + k] == ((c >> 32) == 0 ? uint32_t(c) : ids[i]);
        }
    EXPECT_EQ(same, true);
    std::vector<uint64_t> cross;
    H::ExpandCrossSegment(0, ids, cross);
    EXPECT_EQ(cross.size(), (size_t)0);  // disk 17 cannot move in 12 moves
}

TEST_GPU(HanoiSearch17) {
    // HanoiTowersTests/HanoiSearchTests.cpp:15-23
    HanoiTowers<17> towers;
    std::string initial = towers.ToString(0);
    SearchOptions opts;
    opts.maxSteps = 120;
    auto result = HanoiSearch<towers.Size>(initial, opts);
    EXPECT_EQ(Join(result), std::string(HANOI17_120));
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
