// HanoiTowers<size>, GpuExpandInSegment<size> and HanoiSearch<size> over the C
// ABI (HanoiTowers/HanoiTowers.cpp, HanoiExpandGPU.cu, HanoiSearch.cpp).
#include "HanoiTowers.h"

#include <iostream>
#include <mutex>
#include <sstream>

#include "GPU.h"
#include "HanoiSearch.h"
#include "Util.h"
#include "puzzlesearch_amd.h"

namespace {

// GpuBuffer of HanoiTowers/HanoiExpandGPU.h:7-17: a stream and 16 Mi-entry
// index and child buffers, pooled across calls
struct HanoiGpuBuffer {
    static constexpr size_t SRCBUFSIZE = 16 * 1024 * 1024;
    static constexpr size_t DSTBUFSIZE = 16 * 1024 * 1024;
    CuStream Stream;
    uint32_t* Indexes;
    uint32_t* Children;
    HanoiGpuBuffer()
        : Stream(CreateCudaStream()), Indexes(CreateGPUBuffer32(SRCBUFSIZE)), Children(CreateGPUBuffer32(DSTBUFSIZE)) {}
    ~HanoiGpuBuffer() {
        DestroyCudaStream(Stream);
        DestroyGPUBuffer32(Indexes);
        DestroyGPUBuffer32(Children);
    }
};

std::mutex g_pool_mutex;
std::vector<HanoiGpuBuffer*>& Pool() {
    static std::vector<HanoiGpuBuffer*> pool;
    return pool;
}

HanoiGpuBuffer* Acquire() {
    {
        std::lock_guard<std::mutex> g(g_pool_mutex);
        if (!Pool().empty()) {
            auto* b = Pool().back();
            Pool().pop_back();
            return b;
        }
    }
    return new HanoiGpuBuffer();
}

void Release(HanoiGpuBuffer* b) {
    std::lock_guard<std::mutex> g(g_pool_mutex);
    Pool().push_back(b);
}

void ExpandSlots(int size, int segment, size_t count, const uint32_t* indexes, std::vector<uint32_t>& children,
                 bool withoutSmallest) {
    const size_t dstCount = count * (withoutSmallest ? 3 : 6);
    ensure(count <= HanoiGpuBuffer::SRCBUFSIZE);
    ensure(dstCount <= HanoiGpuBuffer::DSTBUFSIZE);
    children.resize(dstCount);
    if (count == 0) return;
    auto* gpu = Acquire();
    try {
        CopyToGpu(indexes, gpu->Indexes, count, gpu->Stream);
        ERR(ps_hanoi_expand_in_segment(size, (uint32_t)segment, gpu->Indexes, count, gpu->Children,
                                       withoutSmallest ? 1 : 0, gpu->Stream));
        CopyFromGpu(gpu->Children, &children[0], dstCount, gpu->Stream);
    } catch (...) {
        Release(gpu);
        throw;
    }
    Release(gpu);
}

// HanoiTowers::Expand on the device, six children per index
void ExpandFull(int size, const uint64_t* indexes, size_t count, std::vector<uint64_t>& out) {
    out.resize(count * 6);
    if (count == 0) return;
    EnsurePsRuntime();
    void* din = nullptr;
    void* dout = nullptr;
    ERR(ps_alloc(&din, count * 8));
    if (int rc = ps_alloc(&dout, count * 48)) {
        ps_free(din);
        ERR(rc);
    }
    ps_stream s = nullptr;
    int rc = ps_memcpy_h2d_async(din, indexes, count * 8, s);
    if (rc == PS_OK) rc = ps_hanoi_expand(size, (const uint64_t*)din, count, (uint64_t*)dout, s);
    if (rc == PS_OK) rc = ps_memcpy_d2h_async(out.data(), dout, count * 48, s);
    if (rc == PS_OK) rc = ps_stream_sync(s);
    ps_free(din);
    ps_free(dout);
    ERR(rc);
}

}  // namespace

template <int size>
std::string HanoiTowers<size>::ToString(uint64_t index) {
    std::ostringstream stream;
    for (int peg = 0; peg < 4; peg++) {
        for (int disk = 0; disk < size; disk++)
            if (int((index >> (2 * disk)) & 3) == peg) stream << (disk + 1) << ' ';
        stream << 0 << ' ';
    }
    const std::string str = stream.str();
    return str.substr(0, str.size() - 3);
}

template <int size>
uint64_t HanoiTowers<size>::Parse(std::string stateStr) {
    std::istringstream stream(stateStr);
    uint64_t index = 0;
    int peg = 0, value = 0;
    while (stream >> value) {
        if (value == 0) {
            peg++;
        } else {
            ensure(value >= 1 && value <= size && peg < 4);
            index |= uint64_t(peg) << (2 * (value - 1));
        }
    }
    return index;
}

template <int size>
void HanoiTowers<size>::Expand(const std::vector<uint64_t>& indexes, std::vector<uint64_t>& children) {
    std::vector<uint64_t> out;
    ExpandFull(size, indexes.data(), indexes.size(), out);
    children.insert(children.end(), out.begin(), out.end());
}

template <int size>
void HanoiTowers<size>::ExpandInSegment(int segment, size_t count, const uint32_t* indexes,
                                        std::vector<uint32_t>& children) {
    GpuExpandInSegment<size>(segment, count, indexes, children);
}

template <int size>
void HanoiTowers<size>::ExpandInSegmentWithoutSmallest(int segment, size_t count, const uint32_t* indexes,
                                                       std::vector<uint32_t>& children) {
    GpuExpandInSegmentWithoutSmallest<size>(segment, count, indexes, children);
}

template <int size>
void HanoiTowers<size>::ExpandCrossSegment(int segment, const std::vector<uint32_t>& indexes,
                                           std::vector<uint64_t>& children) {
    std::vector<uint64_t> full(indexes.size());
    const uint64_t base = uint64_t(segment) << 32;
    for (size_t i = 0; i < indexes.size(); i++) full[i] = base | indexes[i];
    std::vector<uint64_t> out;
    ExpandFull(size, full.data(), full.size(), out);
    for (uint64_t child : out)
        if (int(child >> 32) != segment) children.push_back(child);
}

template <int size>
void GpuExpandInSegment(int segment, size_t count, const uint32_t* indexes, std::vector<uint32_t>& children) {
    ExpandSlots(size, segment, count, indexes, children, false);
}

template <int size>
void GpuExpandInSegmentWithoutSmallest(int segment, size_t count, const uint32_t* indexes,
                                       std::vector<uint32_t>& children) {
    ExpandSlots(size, segment, count, indexes, children, true);
}

template <int size>
std::vector<uint64_t> HanoiSearch(std::string initialState, SearchOptions options) {
    Timer timer;
    std::cerr << "HanoiSearch (three-bit BFS, segments in HBM)" << std::endl;
    EnsurePsRuntime();
    const uint64_t start = HanoiTowers<size>::Parse(initialState);
    ps_hanoi_search* h = nullptr;
    ERR(ps_hanoi_search_create(size, 32, 0, nullptr, &h));
    const int maxLayers = options.maxSteps < 0 ? 1 : (options.maxSteps >= (1 << 20) ? (1 << 20) : options.maxSteps + 1);
    std::vector<uint64_t> restored(maxLayers), reduced(maxLayers);
    int32_t layers = 0;
    const int rc = ps_hanoi_search_run(h, start, maxLayers, restored.data(), reduced.data(), nullptr, maxLayers, &layers);
    if (rc != PS_OK) {
        const std::string msg = ps_last_error();
        ps_hanoi_search_destroy(h);
        throw std::runtime_error(msg);
    }
    ps_hanoi_search_stats st{};
    ps_hanoi_search_stats_get(h, &st);
    ps_hanoi_search_destroy(h);
    restored.resize(layers);
    uint64_t sum = 0;
    for (auto w : restored) sum += w;
    std::cerr << "Radius: " << restored.size() - 1 << "\nTotal: " << sum << "\nSegments in HBM: " << st.segments_allocated
              << "\nTime: " << timer.Seconds() << " s" << std::endl;
    return restored;
}

#define PS_HANOI_INSTANTIATE(n)                                                                                  \
    template class HanoiTowers<n>;                                                                               \
    template void GpuExpandInSegment<n>(int, size_t, const uint32_t*, std::vector<uint32_t>&);                  \
    template void GpuExpandInSegmentWithoutSmallest<n>(int, size_t, const uint32_t*, std::vector<uint32_t>&);   \
    template std::vector<uint64_t> HanoiSearch<n>(std::string, SearchOptions);

PS_HANOI_INSTANTIATE(14)
PS_HANOI_INSTANTIATE(15)
PS_HANOI_INSTANTIATE(16)
PS_HANOI_INSTANTIATE(17)
PS_HANOI_INSTANTIATE(18)
PS_HANOI_INSTANTIATE(19)
PS_HANOI_INSTANTIATE(20)
PS_HANOI_INSTANTIATE(21)
PS_HANOI_INSTANTIATE(22)
PS_HANOI_INSTANTIATE(23)
PS_HANOI_INSTANTIATE(24)
