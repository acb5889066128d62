// A CPU emulation of the parts of HIP that the BFS kernels of
// puzzlesearch_amd/csrc use, for running those kernels, unchanged, on the
// host in tests:
//   * a block is one OS thread; its lanes are fibers (ucontext) scheduled
//     cooperatively, switching only at __syncthreads and at wave operations
//     (__shfl*, __ballot, wave barriers), so the code between two of them runs
//     in one lane at a time, as a 64-lane wave would run it lane by lane;
//   * __shared__ variables are thread_local: one copy per block (OS thread);
//     the dynamic LDS array is a thread_local buffer whose bytes beyond the
//     launch's request are poisoned for AddressSanitizer;
//   * several blocks of a grid run concurrently on OS threads, so the device
//     atomics (std::atomic_ref here) are exercised across blocks;
//   * every barrier and wave operation is checked: a block barrier must be
//     reached by every lane that has not finished, a wave operation by all 64
//     lanes of the wave with the same operation; a block in which no lane can
//     proceed is reported as a deadlock, and a launch that runs past a time
//     limit is reported as a hang, with the kernel name and block.
// Violations abort with a message (the test reads it).  This is test
// infrastructure: nothing in the library includes it.
#pragma once

#include <sanitizer/asan_interface.h>
#include <ucontext.h>

#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
// block-scope thread_local is static storage: one copy per OS thread = per block
#define __shared__ thread_local

#define __ATOMIC_RELAXED_HIP 0
#define __HIP_MEMORY_SCOPE_WAVEFRONT 1
#define __HIP_MEMORY_SCOPE_WORKGROUP 2
#define __HIP_MEMORY_SCOPE_AGENT 3
#define __HIP_MEMORY_SCOPE_SYSTEM 4

struct dim3 {
    uint32_t x = 1, y = 1, z = 1;
    dim3() = default;
    dim3(uint32_t a, uint32_t b = 1, uint32_t c = 1) : x(a), y(b), z(c) {}
};
struct uint4 {
    uint32_t x, y, z, w;
};
struct ulonglong2 {
    unsigned long long x, y;
};
struct uint2 {
    uint32_t x, y;
};
inline uint4 make_uint4(uint32_t a, uint32_t b, uint32_t c, uint32_t d) { return uint4{a, b, c, d}; }
inline ulonglong2 make_ulonglong2(unsigned long long a, unsigned long long b) { return ulonglong2{a, b}; }
inline uint2 make_uint2(uint32_t a, uint32_t b) { return uint2{a, b}; }

// stand-ins for the runtime types the headers name in host declarations
typedef int hipError_t;
typedef void* hipStream_t;
typedef void* hipEvent_t;
constexpr int hipSuccess = 0;

namespace simt {

enum Op : int { RUN = 0, BAR, WSYNC, SHFL, SHFL_UP, SHFL_XOR, BALLOT, DONE };

struct Lane {
    ucontext_t ctx;
    int tid = 0;
    Op op = RUN;
    uint64_t in = 0, arg = 0, out = 0;
    char* stack = nullptr;
};

// lane stacks, allocated once per block thread and reused by its blocks
constexpr size_t STACK_BYTES = 256 * 1024;
struct StackPool {
    std::vector<char*> s;
    char* get(size_t i) {
        while (s.size() <= i) s.push_back(static_cast<char*>(std::malloc(STACK_BYTES)));
        return s[i];
    }
    ~StackPool() {
        for (char* p : s) std::free(p);
    }
};
inline thread_local StackPool t_stacks;

[[noreturn]] inline void die(const char* fmt, const char* a = "", long b = 0, long c = 0) {
    std::fprintf(stderr, "SIMT: ");
    std::fprintf(stderr, fmt, a, b, c);
    std::fprintf(stderr, "\n");
    std::fflush(stderr);
    std::abort();
}

struct Block {
    ucontext_t sched;
    std::vector<Lane> lanes;
    int tpb = 0;
    dim3 bidx, gdim, bdim;
    const char* kernel = "";
    std::function<void()> body;
    std::atomic<uint64_t>* progress = nullptr;
    long barriers = 0, waveops = 0;
};

inline thread_local Block* t_blk = nullptr;
inline thread_local Lane* t_lane = nullptr;

// dynamic LDS: the kernels' `extern __shared__` array, a thread_local array of
// LDS_BYTES defined by the including test (one per block thread); lds() gives
// the calling thread's copy
constexpr size_t LDS_BYTES = 160 * 1024;
inline unsigned char* (*lds)() = nullptr;

inline void lane_entry() {
    Block* b = t_blk;
    b->body();
    t_lane->op = DONE;
    // back to the scheduler for good
    swapcontext(&t_lane->ctx, &b->sched);
}

inline uint64_t yield_op(Op op, uint64_t in = 0, uint64_t arg = 0) {
    Lane* l = t_lane;
    l->op = op;
    l->in = in;
    l->arg = arg;
    swapcontext(&l->ctx, &t_blk->sched);
    return l->out;
}

// Resolves the wave operations and the block barrier once every lane waits.
inline bool release(Block* b) {
    bool released = false;
    const int nw = (b->tpb + 63) / 64;
    for (int w = 0; w < nw; ++w) {
        Lane* L = &b->lanes[w * 64];
        const int n = std::min(64, b->tpb - w * 64);
        const Op op = L[0].op;
        if (op < WSYNC || op > BALLOT) {
            for (int i = 1; i < n; ++i)
                if (L[i].op >= WSYNC && L[i].op <= BALLOT)
                    die("%s: wave op of lane %ld with lane 0 of its wave not at one (block %ld)", b->kernel,
                        (long)L[i].tid, (long)b->bidx.x);
            continue;
        }
        for (int i = 1; i < n; ++i)
            if (L[i].op != op)
                die("%s: lanes of one wave at different wave operations (lane %ld, block %ld)", b->kernel,
                    (long)L[i].tid, (long)b->bidx.x);
        switch (op) {
        case WSYNC:
            break;
        case BALLOT: {
            uint64_t m = 0;
            for (int i = 0; i < n; ++i) m |= (L[i].in ? 1ull : 0ull) << i;
            for (int i = 0; i < n; ++i) L[i].out = m;
            break;
        }
        case SHFL:
            for (int i = 0; i < n; ++i) L[i].out = L[(int)(L[i].arg % 64) % n].in;
            break;
        case SHFL_UP:
            for (int i = 0; i < n; ++i) L[i].out = (i >= (int)L[i].arg) ? L[i - (int)L[i].arg].in : L[i].in;
            break;
        case SHFL_XOR:
            for (int i = 0; i < n; ++i) L[i].out = L[(i ^ (int)L[i].arg) % n].in;
            break;
        default:
            break;
        }
        for (int i = 0; i < n; ++i) L[i].op = RUN;
        b->waveops++;
        released = true;
    }
    if (released) return true;
    // the block barrier: every lane that has not finished waits at it
    int at_bar = 0, done = 0;
    for (auto& l : b->lanes) {
        if (l.op == BAR) ++at_bar;
        else if (l.op == DONE) ++done;
    }
    if (at_bar > 0 && at_bar + done == b->tpb) {
        if (done) die("%s: %ld lanes finished while the others wait at a block barrier (block %ld)", b->kernel, done,
                      (long)b->bidx.x);
        for (auto& l : b->lanes) l.op = RUN;
        b->barriers++;
        return true;
    }
    return false;
}

inline void run_block(Block* b) {
    t_blk = b;
    b->lanes.resize(b->tpb);
    for (int i = 0; i < b->tpb; ++i) {
        Lane& l = b->lanes[i];
        l.tid = i;
        l.op = RUN;
        l.stack = t_stacks.get((size_t)i);
        getcontext(&l.ctx);
        l.ctx.uc_stack.ss_sp = l.stack;
        l.ctx.uc_stack.ss_size = STACK_BYTES;
        l.ctx.uc_link = nullptr;
        makecontext(&l.ctx, (void (*)())lane_entry, 0);
    }
    for (;;) {
        bool ran = false;
        for (auto& l : b->lanes) {
            if (l.op != RUN) continue;
            t_lane = &l;
            swapcontext(&b->sched, &l.ctx);
            ran = true;
            if (b->progress) b->progress->fetch_add(1, std::memory_order_relaxed);
        }
        bool all_done = true;
        for (auto& l : b->lanes) all_done &= l.op == DONE;
        if (all_done) break;
        if (!release(b) && !ran) {
            int ops[8] = {};
            for (auto& l : b->lanes) ops[l.op]++;
            std::fprintf(stderr, "SIMT: %s: deadlock in block %u: run %d bar %d wsync %d shfl %d/%d/%d ballot %d done %d\n",
                         b->kernel, b->bidx.x, ops[0], ops[1], ops[2], ops[3], ops[4], ops[5], ops[6], ops[7]);
            std::abort();
        }
    }
    t_lane = nullptr;
    t_blk = nullptr;
}

struct LaunchStats {
    long blocks = 0, barriers = 0, waveops = 0;
};

// Runs body() as grid x tpb lanes; `parallel` blocks at a time on OS threads.
// lds_bytes beyond the request are poisoned.  time_limit_s bounds the launch.
inline LaunchStats launch(const char* kernel, uint32_t grid, int tpb, size_t lds_bytes, const std::function<void()>& body,
                          int parallel = 4, double time_limit_s = 120.0) {
    if (lds_bytes > LDS_BYTES) die("%s: launch asks %ld bytes of LDS (emulator holds %ld)", kernel, (long)lds_bytes,
                                   (long)LDS_BYTES);
    if (tpb < 1 || tpb > 1024) die("%s: %ld threads per block", kernel, (long)tpb);
    std::atomic<uint32_t> next{0};
    std::atomic<uint64_t> progress{0};
    std::atomic<bool> finished{false};
    std::mutex mu;
    LaunchStats total;
    auto worker = [&]() {
        unsigned char* t_lds = lds ? lds() : nullptr;
        if (t_lds) ASAN_POISON_MEMORY_REGION(t_lds + lds_bytes, LDS_BYTES - lds_bytes);
        for (;;) {
            const uint32_t bi = next.fetch_add(1);
            if (bi >= grid) break;
            // a fresh block starts with garbage LDS, as the device gives it
            if (t_lds) std::memset(t_lds, 0xA5, lds_bytes);
            Block b;
            b.tpb = tpb;
            b.bidx = dim3(bi);
            b.gdim = dim3(grid);
            b.bdim = dim3((uint32_t)tpb);
            b.kernel = kernel;
            b.body = body;
            b.progress = &progress;
            run_block(&b);
            std::lock_guard<std::mutex> g(mu);
            total.blocks++;
            total.barriers += b.barriers;
            total.waveops += b.waveops;
        }
        if (t_lds) ASAN_UNPOISON_MEMORY_REGION(t_lds, LDS_BYTES);
    };
    std::thread watchdog([&]() {
        const auto t0 = std::chrono::steady_clock::now();
        while (!finished.load()) {
            std::this_thread::sleep_for(std::chrono::milliseconds(50));
            if (std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > time_limit_s) {
                std::fprintf(stderr, "SIMT: %s: launch of %u blocks still running after %.0f s (hang)\n", kernel, grid,
                             time_limit_s);
                std::fflush(stderr);
                std::abort();
            }
        }
    });
    std::vector<std::thread> th;
    const int nth = (int)std::max<uint32_t>(1, std::min<uint32_t>((uint32_t)parallel, grid));
    for (int i = 0; i < nth; ++i) th.emplace_back(worker);
    for (auto& t : th) t.join();
    finished = true;
    watchdog.join();
    return total;
}

template <class T>
inline T bits_as(uint64_t v) {
    T t;
    std::memcpy(&t, &v, sizeof(T));
    return t;
}
template <class T>
inline uint64_t as_bits(T t) {
    uint64_t v = 0;
    std::memcpy(&v, &t, sizeof(T));
    return v;
}

}  // namespace simt

// ---------------------------------------------------------------- the HIP names
#define threadIdx (::simt::t_lane_dim())
#define blockIdx (::simt::t_blk->bidx)
#define gridDim (::simt::t_blk->gdim)
#define blockDim (::simt::t_blk->bdim)
namespace simt {
inline dim3 t_lane_dim() { return dim3((uint32_t)t_lane->tid); }
}  // namespace simt

inline void __syncthreads() { simt::yield_op(simt::BAR); }

#define __builtin_amdgcn_wave_barrier() (::simt::yield_op(::simt::WSYNC))
#define __builtin_amdgcn_fence(order, scope) ((void)0)

template <class T>
inline T __shfl(T v, int src, int width = 64) {
    (void)width;
    return simt::bits_as<T>(simt::yield_op(simt::SHFL, simt::as_bits(v), (uint64_t)src));
}
template <class T>
inline T __shfl_up(T v, unsigned d, int width = 64) {
    (void)width;
    return simt::bits_as<T>(simt::yield_op(simt::SHFL_UP, simt::as_bits(v), d));
}
template <class T>
inline T __shfl_xor(T v, int m, int width = 64) {
    (void)width;
    return simt::bits_as<T>(simt::yield_op(simt::SHFL_XOR, simt::as_bits(v), (uint64_t)m));
}
inline unsigned long long __ballot(int pred) { return simt::yield_op(simt::BALLOT, pred ? 1 : 0); }

inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline int __popc(unsigned int x) { return __builtin_popcount(x); }
inline int __ffsll(long long x) { return __builtin_ffsll(x); }
inline int __ffsll(unsigned long long x) { return __builtin_ffsll((long long)x); }
inline int __ffs(int x) { return __builtin_ffs(x); }
inline int __clz(int x) { return x ? __builtin_clz((unsigned)x) : 32; }
inline int __clzll(long long x) { return x ? __builtin_clzll((unsigned long long)x) : 64; }
inline unsigned int __umul24(unsigned int a, unsigned int b) { return (a & 0xffffffu) * (b & 0xffffffu); }

template <class T>
inline T atomicAdd(T* p, T v) {
    return std::atomic_ref<T>(*p).fetch_add(v);
}
template <class T>
inline T atomicOr(T* p, T v) {
    return std::atomic_ref<T>(*p).fetch_or(v);
}
template <class T>
inline T atomicAnd(T* p, T v) {
    return std::atomic_ref<T>(*p).fetch_and(v);
}
template <class T>
inline T atomicXor(T* p, T v) {
    return std::atomic_ref<T>(*p).fetch_xor(v);
}
template <class T>
inline T atomicCAS(T* p, T cmp, T v) {
    std::atomic_ref<T>(*p).compare_exchange_strong(cmp, v);
    return cmp;
}
template <class T>
inline T atomicMax(T* p, T v) {
    std::atomic_ref<T> a(*p);
    T cur = a.load();
    while (cur < v && !a.compare_exchange_weak(cur, v)) {
    }
    return cur;
}
template <class T, class U>
inline T __hip_atomic_fetch_or(T* p, U v, int, int) {
    return std::atomic_ref<T>(*p).fetch_or((T)v);
}
template <class T, class U>
inline T __hip_atomic_fetch_add(T* p, U v, int, int) {
    return std::atomic_ref<T>(*p).fetch_add((T)v);
}
template <class T>
inline T __hip_atomic_load(const T* p, int, int) {
    return std::atomic_ref<T>(*const_cast<T*>(p)).load();
}
template <class T, class U>
inline void __hip_atomic_store(T* p, U v, int, int) {
    std::atomic_ref<T>(*p).store((T)v);
}
