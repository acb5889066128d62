// GPU puzzle plugins: host-side text formats and ranks, device expansion
// through the launchers of PancakeGpu.h, SlidingPuzzleGpu.h and
// HanoiTowersGPU.h.
#include "Puzzles.h"

#include <algorithm>
#include <sstream>

#include "HanoiTowersGPU.h"
#include "PancakeGpu.h"
#include "SlidingPuzzleGpu.h"

namespace ps_host {

// digit_i = number of values smaller than a[i] that do not occur before i
uint64_t PermutationRank(const int* arr, int size) {
    uint64_t rank = 0;
    uint32_t used = 0;
    for (int i = 0; i < size; i++) {
        const uint32_t below = (1u << arr[i]) - 1u;
        rank = rank * (uint64_t)(size - i) + (uint64_t)__builtin_popcount(below & ~used);
        used |= 1u << arr[i];
    }
    return rank;
}

void PermutationUnrank(uint64_t index, int* arr, int size) {
    int digits[32];
    for (int i = size - 1; i >= 0; i--) {
        digits[i] = (int)(index % (uint64_t)(size - i));
        index /= (uint64_t)(size - i);
    }
    uint32_t used = 0;
    for (int i = 0; i < size; i++) {
        int k = digits[i], v = 0;
        for (;; v++) {
            if (used & (1u << v)) continue;
            if (k-- == 0) break;
        }
        arr[i] = v;
        used |= 1u << v;
    }
}

}  // namespace ps_host

namespace {

std::vector<int> ReadInts(const std::string& s) {
    std::istringstream in(s);
    std::vector<int> v;
    int x;
    while (in >> x) v.push_back(x);
    return v;
}

std::string JoinInts(const int* a, int n) {
    std::ostringstream out;
    for (int i = 0; i < n; i++) {
        if (i) out << ' ';
        out << a[i];
    }
    return out.str();
}

uint64_t Factorial(int n) {
    uint64_t r = 1;
    for (int i = 2; i <= n; i++) r *= (uint64_t)i;
    return r;
}

// PancakeOptimized's segmented rank (Puzzles/PancakeOptimized.cpp:26-77):
// n <= 12 plain rank; else (rank of the first n-12 digits) << 29 | rank of
// the last 12 digits.
uint64_t PancakeRank(const int* arr, int n) {
    if (n <= 12) return ps_host::PermutationRank(arr, n);
    uint32_t used = 0;
    int d[16];
    for (int i = 0; i < n; i++) {
        d[i] = __builtin_popcount(((1u << arr[i]) - 1u) & ~used);
        used |= 1u << arr[i];
    }
    uint64_t low = 0, seg = 0;
    for (int i = n - 12; i < n; i++) low = low * (uint64_t)(n - i) + (uint64_t)d[i];
    for (int i = 0; i < n - 12; i++) seg = seg * (uint64_t)(n - i) + (uint64_t)d[i];
    return (seg << 29) | low;
}

void PancakeUnrank(uint64_t index, int* arr, int n) {
    if (n <= 12) {
        ps_host::PermutationUnrank(index, arr, n);
        return;
    }
    int d[16];
    uint64_t seg = index >> 29, low = index & ((1ull << 29) - 1);
    for (int i = n - 1; i >= n - 12; i--) {
        d[i] = (int)(low % (uint64_t)(n - i));
        low /= (uint64_t)(n - i);
    }
    for (int i = n - 13; i >= 0; i--) {
        d[i] = (int)(seg % (uint64_t)(n - i));
        seg /= (uint64_t)(n - i);
    }
    uint32_t used = 0;
    for (int i = 0; i < n; i++) {
        int k = d[i], v = 0;
        for (;; v++) {
            if (used & (1u << v)) continue;
            if (k-- == 0) break;
        }
        arr[i] = v;
        used |= 1u << v;
    }
}

}  // namespace

// ---------------------------------------------------------------- pancake

PancakeOptimizedGPU::PancakeOptimizedGPU(int size, bool invIndex) : m_Size(size), m_InvIndex(invIndex) {
    ensure(size > 1 && size <= 16);
}

std::string PancakeOptimizedGPU::Name() const {
    std::ostringstream s;
    s << "Pancake (GPU, 29 bits optimized), size: " << m_Size << "; invIdx: " << m_InvIndex;
    return s.str();
}

uint64_t PancakeOptimizedGPU::IndexesCount() const {
    ps_puzzle p = Descriptor();
    ps_puzzle_info info;
    PsCheck(ps_puzzle_query(&p, &info), "ps_puzzle_query");
    return info.indexes_count;
}

std::string PancakeOptimizedGPU::ToString(uint64_t index) {
    int a[16];
    PancakeUnrank(index, a, m_Size);
    return JoinInts(a, m_Size);
}

uint64_t PancakeOptimizedGPU::Parse(std::string state) {
    auto v = ReadInts(state);
    ensure((int)v.size() >= m_Size);
    return PancakeRank(v.data(), m_Size);
}

ps_puzzle PancakeOptimizedGPU::Descriptor() const {
    return ps_puzzle{PS_PANCAKE, m_Size, 0, m_InvIndex ? PS_FLAG_INV_INDEX : 0};
}

// The reference's body is ensure(false): its Expand calls the launchers
// itself.  Here it expands with the full flip set, like PancakeExpand.
void PancakeOptimizedGPU::ExpandGpu(uint64_t* gpuIndexes, uint64_t* gpuExpanded, uint64_t count, CuStream stream) {
    PancakeExpand(gpuIndexes, gpuExpanded, m_Size, count, stream);
}

// Puzzles/PancakeOptimizedGpu.cpp:36-79: with 29-bit segments and n > 12 the
// hint selects the in-segment (ops 0..10) or cross-segment (ops 11..n-2) slots.
void PancakeOptimizedGPU::Expand(std::vector<uint64_t>& indexes, std::vector<int>& usedOperatorBits,
                                 std::vector<uint64_t>& expandedIndexes, std::vector<int>& expandedOperators,
                                 ExpandHint hint) {
    SetupOutputBuffers(expandedIndexes, expandedOperators);
    if (indexes.empty()) return;
    CombineIndexAndOpBits(indexes, usedOperatorBits);
    const int br = BranchingFactor();
    const int n = m_Size;
    const bool inv = m_InvIndex;
    if (hint.SegmentBits == 29 && n > 12) {
        if (hint.CrossSegment)
            RunBatch(indexes, expandedIndexes, expandedOperators, br - 11,
                     [&](uint64_t* src, uint64_t* dst, uint64_t count, CuStream s) {
                         PancakeExpandCrossSegment(src, dst, n, inv, count, s);
                     });
        else
            RunBatch(indexes, expandedIndexes, expandedOperators, 11,
                     [&](uint64_t* src, uint64_t* dst, uint64_t count, CuStream s) {
                         PancakeExpandInSegment(src, dst, n, count, s);
                     });
    } else {
        RunBatch(indexes, expandedIndexes, expandedOperators, br,
                 [&](uint64_t* src, uint64_t* dst, uint64_t count, CuStream s) { PancakeExpand(src, dst, n, count, s); });
    }
}

// Puzzles/PancakeOptimizedGpu.cpp:21-34
void PancakeOptimizedGPU::CrossSegmentPostProcess(int op, int segment, int segmentBits,
                                                  Buffer<uint32_t>& expandedIndexes) {
    (void)op;
    const bool invert = m_InvIndex && segmentBits == 29 && m_Size > 12;
    if (!invert || expandedIndexes.IsEmpty()) return;
    // the u32 ids are staged in the input buffer of MAX_INDEXES_BUFFER u64
    ensure(expandedIndexes.Size() <= 2 * MAX_INDEXES_BUFFER);
    Exec* ex = AquireStream();
    try {
        uint32_t* gpu = (uint32_t*)ex->gpuSrc;
        CopyToGpu(expandedIndexes.Buf(), gpu, expandedIndexes.Size(), ex->stream);
        PancakeCrossSegmentPostProcessGPU(gpu, segment, m_Size, expandedIndexes.Size(), ex->stream);
        CopyFromGpu(gpu, expandedIndexes.Buf(), expandedIndexes.Size(), ex->stream);
    } catch (...) {
        ReleaseStream(ex);
        throw;
    }
    ReleaseStream(ex);
}

// ---------------------------------------------------------- sliding simple

SlidingTilePuzzleGpu::SlidingTilePuzzleGpu(int width, int height)
    : m_Width(width), m_Height(height), m_Size(width * height) {
    ensure(width > 0 && height > 0 && m_Size >= 2 && m_Size <= 16);
}

std::string SlidingTilePuzzleGpu::Name() const {
    std::ostringstream s;
    s << "Sliding-Tile GPU: " << m_Width << " x " << m_Height;
    return s.str();
}

uint64_t SlidingTilePuzzleGpu::IndexesCount() const { return Factorial(m_Size); }

std::string SlidingTilePuzzleGpu::ToString(uint64_t index) {
    int a[16];
    ps_host::PermutationUnrank(index, a, m_Size);
    return JoinInts(a, m_Size);
}

uint64_t SlidingTilePuzzleGpu::Parse(std::string state) {
    auto v = ReadInts(state);
    ensure((int)v.size() >= m_Size);
    return ps_host::PermutationRank(v.data(), m_Size);
}

ps_puzzle SlidingTilePuzzleGpu::Descriptor() const { return ps_puzzle{PS_SLIDING, m_Width, m_Height, 0}; }

// Puzzles/SlidingTilePuzzleGpu.cpp:49-56
void SlidingTilePuzzleGpu::ExpandGpu(uint64_t* gpuIndexes, uint64_t* gpuExpanded, uint64_t count, CuStream stream) {
    GpuSlidingTilePuzzleSimpleExpand(gpuIndexes, gpuExpanded, m_Width, m_Size, count, stream);
}

// ---------------------------------------------------------- sliding packed

namespace {

// Non-blank tiles (tile-1) in cell order, the blank cell in t[15]; see
// Puzzles/SlidingTilePuzzleOptimized.cpp:79-192 for the packed layout.
bool PackedParityOk(int width, int size, const int* t) {
    int inversions = 0;
    for (int i = 0; i < size - 1; i++)
        for (int j = 0; j < i; j++) inversions += t[j] > t[i];
    const int row = t[15] / width;
    return (width % 2 == 0) ? ((row + inversions) % 2 == 0) : (inversions % 2 == 0);
}

uint64_t PackedIndexOfCells(int width, int height, const std::vector<int>& cells) {
    const int size = width * height;
    int t[16] = {0};
    int pos = 0, blank = -1;
    bool seen[16] = {false};
    for (int i = 0; i < size; i++) {
        ensure(cells[i] >= 0 && cells[i] < size && !seen[cells[i]]);
        seen[cells[i]] = true;
        if (cells[i] == 0) blank = i;
        else t[pos++] = cells[i] - 1;
    }
    t[15] = blank;
    ensure(PackedParityOk(width, size, t));
    int packed[16] = {0};
    for (int i = 2; i < size - 1; i++) {
        int c = 0;
        for (int j = 0; j < i; j++) c += t[j] < t[i];
        packed[i] = c;
    }
    const uint32_t seg = (uint32_t)(packed[14] * 256 + packed[13] * 16 + packed[12]);
    uint32_t idx = 0;
    for (int i = 11; i >= 2; i--) idx = idx * (uint32_t)(i + 1) + (uint32_t)packed[i];
    return ((uint64_t)seg << 32) | (uint64_t)(idx * 16u + (uint32_t)blank);
}

std::vector<int> PackedCellsOfIndex(int width, int height, uint64_t index) {
    const int size = width * height, m = size - 1;
    int packed[16] = {0};
    const uint32_t seg = (uint32_t)(index >> 32);
    uint32_t idx = (uint32_t)index;
    packed[12] = seg % 16;
    packed[13] = (seg / 16) % 16;
    packed[14] = seg / 256;
    const int blank = (int)(idx % 16);
    idx /= 16;
    for (int i = 2; i < 12; i++) {
        packed[i] = (int)(idx % (uint32_t)(i + 1));
        idx /= (uint32_t)(i + 1);
    }
    // decode "smaller tiles to the left" from the right end; position 1 first
    // gets the smaller of the last two values, then parity decides the order
    std::vector<int> remaining;
    for (int v = 0; v < m; v++) remaining.push_back(v);
    int t[16] = {0};
    for (int i = m - 1; i >= 1; i--) {
        const int c = (i >= 2) ? packed[i] : 0;
        ensure(c < (int)remaining.size());
        t[i] = remaining[c];
        remaining.erase(remaining.begin() + c);
    }
    t[0] = remaining[0];
    t[15] = blank;
    if (!PackedParityOk(width, size, t)) std::swap(t[0], t[1]);
    std::vector<int> cells(size);
    int pos = 0;
    for (int i = 0; i < size; i++) cells[i] = (i == blank) ? 0 : t[pos++] + 1;
    return cells;
}

}  // namespace

template <int width, int height>
std::string SlidingTilePuzzleOptimized<width, height>::Name() const {
    std::ostringstream s;
    s << "Sliding-Tile GPU Optimized (for 32-bit segment): " << width << " x " << height;
    return s.str();
}

template <int width, int height>
uint64_t SlidingTilePuzzleOptimized<width, height>::IndexesCount() const {
    ps_puzzle p = Descriptor();
    ps_puzzle_info info;
    PsCheck(ps_puzzle_query(&p, &info), "ps_puzzle_query");
    return info.indexes_count;
}

template <int width, int height>
std::string SlidingTilePuzzleOptimized<width, height>::ToString(uint64_t index) {
    auto cells = PackedCellsOfIndex(width, height, index);
    return JoinInts(cells.data(), size);
}

template <int width, int height>
uint64_t SlidingTilePuzzleOptimized<width, height>::Parse(std::string state) {
    auto v = ReadInts(state);
    ensure((int)v.size() >= size);
    v.resize(size);
    return PackedIndexOfCells(width, height, v);
}

template <int width, int height>
void SlidingTilePuzzleOptimized<width, height>::ExpandGpu(uint64_t* gpuIndexes, uint64_t* gpuExpanded, uint64_t count,
                                                          CuStream stream) {
    GpuSlidingTilePuzzleOptimizedExpand<width, height>(gpuIndexes, gpuExpanded, count, stream);
}

// Puzzles/SlidingTilePuzzleOptimized.cpp:247-285: with 32-bit segments only
// parents whose blank sits where a vertical move can change the segment are
// expanded for the cross-segment pass.
template <int width, int height>
void SlidingTilePuzzleOptimized<width, height>::Expand(std::vector<uint64_t>& indexes,
                                                       std::vector<int>& usedOperatorBits,
                                                       std::vector<uint64_t>& expandedIndexes,
                                                       std::vector<int>& expandedOperators, ExpandHint hint) {
    SetupOutputBuffers(expandedIndexes, expandedOperators);
    if (hint.SegmentBits == 32 && size > 12 && hint.CrossSegment) {
        size_t dst = 0;
        for (size_t i = 0; i < indexes.size(); i++) {
            const int b = (int)(indexes[i] & 15);
            const bool keep = b == 13 || b == 14 || b == 15 || b == 13 - width || b == 14 - width || b == 15 - width;
            if (!keep) continue;
            indexes[dst] = indexes[i];
            usedOperatorBits[dst] = usedOperatorBits[i];
            dst++;
        }
        indexes.resize(dst);
        usedOperatorBits.resize(dst);
    }
    if (indexes.empty()) return;
    CombineIndexAndOpBits(indexes, usedOperatorBits);
    RunBatch(indexes, expandedIndexes, expandedOperators, 4,
             [&](uint64_t* src, uint64_t* dst, uint64_t count, CuStream s) { ExpandGpu(src, dst, count, s); });
}

template class SlidingTilePuzzleOptimized<2, 2>;
template class SlidingTilePuzzleOptimized<3, 2>;
template class SlidingTilePuzzleOptimized<4, 2>;
template class SlidingTilePuzzleOptimized<5, 2>;
template class SlidingTilePuzzleOptimized<6, 2>;
template class SlidingTilePuzzleOptimized<7, 2>;
template class SlidingTilePuzzleOptimized<8, 2>;
template class SlidingTilePuzzleOptimized<3, 3>;
template class SlidingTilePuzzleOptimized<4, 3>;
template class SlidingTilePuzzleOptimized<5, 3>;
template class SlidingTilePuzzleOptimized<4, 4>;

// ------------------------------------------------------------ 4-peg Hanoi

FourPegHanoiGPUBase::FourPegHanoiGPUBase(int size, bool useSymmetry, bool optimized)
    : m_Size(size), m_UseSymmetry(useSymmetry), m_Optimized(optimized) {
    ensure(size >= 1 && size <= 24);
}

std::string FourPegHanoiGPUBase::Name() const {
    std::ostringstream s;
    s << (m_Optimized ? "Four-Peg Hanoi Optimized (GPU), size=" : "Four-Peg Hanoi Towers (GPU), size=") << m_Size
      << "; symmetry=" << m_UseSymmetry;
    return s.str();
}

// Puzzles/FourPegHanoiSimple.cpp:114-124: disks 1..n of each peg from the top,
// pegs separated by 0, trailing " 0" of the last peg dropped.
std::string FourPegHanoiGPUBase::ToString(uint64_t index) {
    std::ostringstream s;
    for (int peg = 0; peg < 4; peg++) {
        for (int d = 0; d < m_Size; d++)
            if ((int)((index >> (2 * d)) & 3) == peg) s << (d + 1) << ' ';
        s << "0 ";
    }
    std::string str = s.str();
    return str.substr(0, str.size() - 3);
}

uint64_t FourPegHanoiGPUBase::Parse(std::string state) {
    uint64_t index = 0;
    int peg = 0;
    for (int v : ReadInts(state)) {
        if (v == 0) peg++;
        else {
            ensure(v >= 1 && v <= m_Size && peg < 4);
            index |= (uint64_t)peg << (2 * (v - 1));
        }
    }
    return index;
}

ps_puzzle FourPegHanoiGPUBase::Descriptor() const {
    return ps_puzzle{m_Optimized ? PS_HANOI4_OPT : PS_HANOI4, m_Size, 0, m_UseSymmetry ? PS_FLAG_SYMMETRY : 0};
}

// Puzzles/FourPegHanoiGPU.cpp:12-25, Puzzles/FourPegHanoiOptimizedGPU.cpp:12-25
void FourPegHanoiGPUBase::ExpandGpu(uint64_t* gpuIndexes, uint64_t* gpuExpanded, uint64_t count, CuStream stream) {
    if (m_Optimized) GpuHanoiTowersExpandOptimized(gpuIndexes, gpuExpanded, m_Size, m_UseSymmetry, count, stream);
    else GpuHanoiTowersExpand(gpuIndexes, gpuExpanded, m_Size, m_UseSymmetry, count, stream);
}

// Puzzles/FourPegHanoiGPU.cpp:27-70: for the cross-segment pass with 32-bit
// segments, parents whose low 16 disks occupy at least three pegs cannot move
// a disk >= 16 and are dropped.
void FourPegHanoiGPUBase::Expand(std::vector<uint64_t>& indexes, std::vector<int>& usedOperatorBits,
                                 std::vector<uint64_t>& expandedIndexes, std::vector<int>& expandedOperators,
                                 ExpandHint hint) {
    if (hint.SegmentBits == 32 && hint.CrossSegment && m_Size > 16) {
        size_t dst = 0;
        for (size_t i = 0; i < indexes.size(); i++) {
            const uint32_t idx = (uint32_t)indexes[i];
            const uint32_t p0 = idx >> 1, p1 = idx, m = 0x55555555u;
            const int pegs = ((p0 & p1 & m) != 0) + ((p0 & ~p1 & m) != 0) + ((~p0 & p1 & m) != 0) +
                             ((~p0 & ~p1 & m) != 0);
            if (pegs >= 3) continue;
            indexes[dst] = indexes[i];
            usedOperatorBits[dst] = usedOperatorBits[i];
            dst++;
        }
        indexes.resize(dst);
        usedOperatorBits.resize(dst);
    }
    PuzzleGpu::Expand(indexes, usedOperatorBits, expandedIndexes, expandedOperators, hint);
}
