// PuzzleGpu (Puzzles/PuzzleGpu.cpp:4-82) over the GPU.h shim.
#include "PuzzleGpu.h"

#include <cstdlib>
#include <string>

#include "GPU.h"

std::atomic<uint64_t> ExpandBuffer::s_nodes{0};
std::atomic<uint64_t> ExpandBuffer::s_nanos{0};
std::atomic<uint64_t> ExpandBuffer::s_times{0};

void ExpandBuffer::PrintStats() {
    std::cerr << "Expand: " << s_times.load() << " times, " << s_nodes.load() << " nodes in "
              << s_nanos.load() / 1e9 << " s" << std::endl;
}

void PsCheck(int status, const char* what) {
    if (status != PS_OK) {
        std::string msg = std::string(what) + ": " + ps_last_error();
        std::cerr << msg << std::endl;
        throw std::runtime_error(msg);
    }
}

// The library is initialised on first use by any GPU plugin (the reference's
// CUDA runtime initialises lazily the same way), unless the program has
// already called ps_init for a device of its choice.
void EnsurePsRuntime() {
    static std::once_flag once;
    static int status = PS_OK;
    std::call_once(once, [] {
        int32_t cus = 0;
        if (ps_device_props(nullptr, 0, &cus, nullptr, nullptr) == PS_OK) return;  // already initialised
        const char* dev = std::getenv("PS_DEVICE");
        status = ps_init(dev ? std::atoi(dev) : 0);
    });
    PsCheck(status, "ps_init");
}

PuzzleGpu::Exec::Exec(int branchingFactor) {
    stream = CreateCudaStream();
    try {
        gpuSrc = CreateGPUBuffer((int)MAX_INDEXES_BUFFER);
        gpuDst = CreateGPUBuffer((int)MAX_INDEXES_BUFFER * branchingFactor);
    } catch (...) {
        if (gpuSrc) ps_free(gpuSrc);
        ps_stream_destroy(stream);
        throw;
    }
}

// Destructors do not throw: failures to release are ignored here.
PuzzleGpu::Exec::~Exec() {
    ps_stream_destroy(stream);
    ps_free(gpuSrc);
    ps_free(gpuDst);
}

PuzzleGpu::Exec* PuzzleGpu::AquireStream() {
    std::lock_guard<std::mutex> lock(m_Mutex);
    if (m_FreeStreams.empty()) {
        m_Streams.emplace_back(std::make_unique<Exec>(BranchingFactor()));
        m_FreeStreams.push_back(m_Streams.back().get());
    }
    Exec* e = m_FreeStreams.back();
    m_FreeStreams.pop_back();
    return e;
}

void PuzzleGpu::ReleaseStream(Exec* stream) {
    std::lock_guard<std::mutex> lock(m_Mutex);
    m_FreeStreams.push_back(stream);
}

// Puzzles/PuzzleGpu.cpp:35-40
void PuzzleGpu::CombineIndexAndOpBits(std::vector<uint64_t>& indexes, std::vector<int>& usedOperatorBits) {
    const int ops = OperatorsCount();
    for (size_t i = 0; i < indexes.size(); i++) indexes[i] = (indexes[i] << ops) | (uint64_t)usedOperatorBits[i];
}

// Puzzles/PuzzleGpu.cpp:42-51: arithmetic shift keeps INVALID_INDEX
void PuzzleGpu::SplitIndexAndOps(std::vector<uint64_t>& expandedIndexes, std::vector<int>& expandedOperators) {
    const int ops = OperatorsCount();
    const uint64_t mask = (1ull << ops) - 1;
    for (size_t i = 0; i < expandedIndexes.size(); i++) {
        const uint64_t v = expandedIndexes[i];
        expandedOperators[i] = (int)(v & mask);
        expandedIndexes[i] = (uint64_t)((int64_t)v >> ops);
    }
}

// Puzzles/PuzzleGpu.cpp:53-82
void PuzzleGpu::Expand(std::vector<uint64_t>& indexes, std::vector<int>& usedOperatorBits,
                       std::vector<uint64_t>& expandedIndexes, std::vector<int>& expandedOperators, ExpandHint hint) {
    (void)hint;
    SetupOutputBuffers(expandedIndexes, expandedOperators);
    if (indexes.empty()) return;
    ensure(indexes.size() <= MAX_INDEXES_BUFFER);
    CombineIndexAndOpBits(indexes, usedOperatorBits);
    RunBatch(indexes, expandedIndexes, expandedOperators, BranchingFactor(),
             [&](uint64_t* src, uint64_t* dst, uint64_t count, CuStream s) { ExpandGpu(src, dst, count, s); });
}
