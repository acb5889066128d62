// In-memory BFS drivers (Common/InMemoryTwoBitBFS.cpp:24-69,
// Common/InMemoryThreeBitBFS.cpp:4-48) and their device-resident variants.
#include "InMemoryBFS.h"

#include <cstring>

#include "BitArray.h"

namespace {

// 2 bits per state: value v of state i at bits 2i..2i+1 of the word array
class TwoBitArray {
public:
    explicit TwoBitArray(uint64_t size) : m_Words((size + 31) / 32, 0) {}
    int Get(uint64_t i) const { return (int)((m_Words[i >> 5] >> (2 * (i & 31))) & 3); }
    void Set(uint64_t i, int v) {
        const uint64_t sh = 2 * (i & 31);
        m_Words[i >> 5] = (m_Words[i >> 5] & ~(3ull << sh)) | ((uint64_t)v << sh);
    }

private:
    std::vector<uint64_t> m_Words;
};

}  // namespace

std::vector<uint64_t> InMemoryTwoBitBFS(Puzzle& puzzle, std::string initialState) {
    const uint64_t SIZE = puzzle.IndexesCount();
    std::vector<uint64_t> result;
    int UNVISITED = 0, OLD = 1, CUR = 2, NEXT = 3;
    ExpandBuffer nodes(puzzle);
    TwoBitArray array(SIZE);
    array.Set(puzzle.Parse(initialState), CUR);
    result.push_back(1);
    while (true) {
        uint64_t count = 0;
        auto fnExpand = [&](uint64_t child, int) {
            if (array.Get(child) == UNVISITED) {
                count++;
                array.Set(child, NEXT);
            }
        };
        for (uint64_t i = 0; i < SIZE; i++) {
            if (array.Get(i) != CUR) continue;
            array.Set(i, OLD);
            nodes.Add(i, 0, fnExpand);
        }
        nodes.Finish(fnExpand);
        if (count == 0) break;
        result.push_back(count);
        std::swap(CUR, NEXT);
    }
    return result;
}

std::vector<uint64_t> InMemoryThreeBitBFS(Puzzle& puzzle, std::string initialState) {
    const uint64_t SIZE = puzzle.IndexesCount();
    std::vector<uint64_t> result;
    BitArray listOld(SIZE), listCur(SIZE), listNew(SIZE);
    listCur.Set(puzzle.Parse(initialState));
    result.push_back(1);
    ExpandBuffer nodes(puzzle);
    auto fnExpand = [&](uint64_t child, int) { listNew.Set(child); };
    while (true) {
        listCur.ScanBits([&](uint64_t index) { nodes.Add(index, 0, fnExpand); });
        nodes.Finish(fnExpand);
        if (puzzle.HasOddLengthCycles()) listNew.AndNot(listCur);
        const uint64_t count = listNew.AndNotAndCount(listOld);
        if (count == 0) break;
        result.push_back(count);
        std::swap(listOld, listCur);
        std::swap(listCur, listNew);
        listNew.Clear();
    }
    return result;
}

namespace {

std::vector<uint64_t> RunDevice(PuzzleGpu& puzzle, const std::string& initialState, const GpuSearchOptions& opts,
                                int bits) {
    const uint64_t start = puzzle.Parse(initialState);
    ps_puzzle p = puzzle.Descriptor();
    ps_bfs_options o;
    ps_bfs_options_default(&o);
    o.bits = bits;
    o.segment_bits = opts.segmentBits;
    o.mode = opts.batched ? PS_BFS_BATCHED : PS_BFS_PLAIN;
    o.max_layers = opts.maxSteps ? (int32_t)(opts.maxSteps + 1) : 0;
    o.index_count = opts.indexCount;
    o.hash = opts.frontierHashes ? 1 : 0;
    EnsurePsRuntime();
    ps_bfs* b = nullptr;
    PsCheck(ps_bfs_create(&p, &o, nullptr, &b), "ps_bfs_create");
    const int cap = 1 << 16;
    std::vector<uint64_t> counts(cap), hashes(cap);
    int32_t n = 0;
    const int rc = ps_bfs_run(b, start, counts.data(), hashes.data(), cap, &n);
    if (rc == PS_OK && opts.stats) ps_bfs_stats_get(b, opts.stats);
    ps_bfs_destroy(b);
    PsCheck(rc, "ps_bfs_run");
    counts.resize(n);
    if (opts.frontierHashes) {
        hashes.resize(n);
        *opts.frontierHashes = hashes;
    }
    return counts;
}

}  // namespace

std::vector<uint64_t> GpuInMemoryThreeBitBFS(PuzzleGpu& puzzle, std::string initialState, const GpuSearchOptions& opts) {
    return RunDevice(puzzle, initialState, opts, 3);
}

std::vector<uint64_t> GpuInMemoryTwoBitBFS(PuzzleGpu& puzzle, std::string initialState, const GpuSearchOptions& opts) {
    return RunDevice(puzzle, initialState, opts, 2);
}
