// Multi-search engine: continuous batching of complete three-bit BFS jobs.
//
// Jobs of different puzzles and sizes are submitted at any time.  A step
//   1 admits queued jobs, first come first served, while their HBM footprint
//     (ps_bfs_footprint) fits the budget and a job slot is free;
//   2 advances the running jobs, in one of the two ways below;
//   3 retires the jobs whose search ended: their layer counts and frontier
//     hashes are copied out; with per-job streams the search itself is kept
//     for the next job of the same puzzle (IdleSearch) and its memory is
//     released only when a queued job needs it, else it is released at once.
// Two ways of running a step:
//   PS_BATCH_STREAMS (default): every job is a ps_bfs on its own stream,
//     advanced by replays of its captured layer graph; each job runs on its
//     own completion event, so a step waits for the first job to finish its
//     layers in flight, not for all of them, and the jobs' kernels overlap on
//     the device (step_streams);
//   PS_BATCH_FUSED: one layer of every running job with three launches shared
//     by all of them: a persistent expand over the concatenated tile lists, a
//     persistent finalize over the concatenated chunk lists and a bookkeeping
//     kernel with one block per job (k_multi_* in bfs_kernels.h).  Disabled: its first
//     run on the device stopped producing output inside a step and no cause has
//     been found, so selecting it fails until one is.
// Each job's arrays, counters and layer index are its own (a ps_bfs), so a job's
// result does not depend on what else is running.  This is the device form of
// running InMemoryThreeBitBFS (Common/InMemoryThreeBitBFS.cpp:4-48) for many
// puzzles at once.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <new>
#include <stdexcept>
#include <thread>
#include <vector>

#include "bfs_internal.h"
#include "multi_kinds.h"
#include "ps_internal.h"

namespace {

constexpr int MULTI_TPB = 512;
constexpr int FIN_TPB = 256;

uint64_t now_ns() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct Job {
    uint64_t id = 0;
    ps_puzzle puzzle{};
    uint64_t start = 0;
    uint64_t index_count = 0;
    uint64_t footprint = 0;
    uint64_t states = 0;  // size of the index space
    int kind = -1;
    int state = PS_JOB_QUEUED;
    ps_bfs* bfs = nullptr;
    uint64_t t_submit = 0, t_admit = 0, t_done = 0;
    uint64_t steps = 0;
    std::vector<uint64_t> counts, hashes;
    uint64_t expanded = 0, children = 0;
    std::string error;
    // per-job streams: layers enqueued and not yet seen complete, and the
    // event recorded after them and the state copy
    hipEvent_t ev = nullptr;
    bool in_flight = false;
    int batch_layers = 2;  // layers enqueued at once (step_streams adapts it)
    uint64_t t_enqueue = 0;
    hipStream_t masked_stream = nullptr;  // the search's stream when it is kept off the reserved CUs
    uint64_t remaining = 0;   // states not yet visited, as of the last completed layers
    uint64_t held_since = 0;  // shortest first: when the job was first held back (0: not held)
};

constexpr int MIN_BATCH_LAYERS = 2, MAX_BATCH_LAYERS = 32;

// A retired job's search (arrays, stream, captured graph) kept for the next job
// of the same puzzle and index space.  Releasing device memory waits for the
// whole device, every other job's layers in flight included, so it is done only
// when a queued job needs the memory or the batch is destroyed.
struct IdleSearch {
    ps_puzzle puzzle{};
    uint64_t index_count = 0;
    uint64_t footprint = 0;
    ps_bfs* bfs = nullptr;
    hipStream_t masked_stream = nullptr;
};

// Searches of at least 2^LATENCY_BITS states are throughput-bound; with
// reserved CUs they run on streams masked to the other CUs, so that the small,
// latency-bound searches (launch chains of short kernels) always find CUs
// free instead of waiting for a large search's persistent kernel to end.
constexpr int LATENCY_BITS = 25;

bool same_search(const ps_puzzle& a, uint64_t na, const ps_puzzle& b, uint64_t nb) {
    return a.kind == b.kind && a.size == b.size && a.height == b.height && a.flags == b.flags && na == nb;
}

}  // namespace

struct ps_batch {
    ps_batch_options opt{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    uint64_t budget = 0, used = 0;
    uint64_t next_id = 1;
    std::map<uint64_t, Job*> jobs;
    std::deque<Job*> queue;
    std::vector<Job*> running;
    std::vector<uint64_t> finished_last;
    psk::MultiJob* d_jobs = nullptr;
    psk::MultiJob* h_jobs = nullptr;  // pinned
    unsigned long long* d_counters = nullptr;
    uint32_t* d_done = nullptr;
    uint32_t* h_done = nullptr;  // pinned
    bool table_dirty = true;
    uint64_t total_tiles = 0, total_chunks = 0;
    uint32_t win_u32_max = 0;
    size_t lds = 0;
    uint32_t grid_expand = 0, grid_fin = 0;
    uint32_t cus = 256;
    ps_batch_stats stats{};
    std::vector<IdleSearch> idle;  // oldest first; their bytes count in `used`
    std::vector<uint32_t> large_mask;  // CU mask of large searches (empty: no CUs reserved)
};

namespace ps {
namespace {

int bfs_opts_for(const ps_batch* b, const Job* j, ps_bfs_options* o) {
    ps_bfs_options_default(o);
    o->bits = 3;
    o->mode = PS_BFS_BATCHED;
    o->graph = 0;
    o->index_count = j->index_count;
    o->block_threads = MULTI_TPB;
    o->cross_marks = PS_CROSS_BYTES;
    ps_puzzle_info info{};
    if (int r = puzzle_query(&j->puzzle, &info)) return r;
    const uint64_t n = j->index_count ? j->index_count : info.indexes_count;
    // per-job streams: 10..12 pancakes run with cross-tile lists, the other
    // searches OR their cross-window children in with atomics, which needs no
    // byte per state (the fused kernel has byte marks only)
    if (b->opt.mode == PS_BATCH_STREAMS) {
        // layers are replayed from a captured graph of MIN_BATCH_LAYERS
        o->graph = 1;
        o->graph_layers = MIN_BATCH_LAYERS;
        o->grid_blocks = b->opt.job_grid_blocks;
        o->cross_marks = PS_CROSS_ATOMIC;
        if (j->puzzle.kind == PS_PANCAKE && j->puzzle.size >= 10 && j->puzzle.size <= 12 &&
            n >= factorial(j->puzzle.size))
            o->cross_marks = PS_CROSS_LISTS;
    }
    // segments pay off once a search is large; small ones run as one segment
    o->segment_bits = (b->opt.segment_bits > 0 && n > (1ull << b->opt.segment_bits)) ? b->opt.segment_bits : 0;
    return PS_OK;
}

// keep_search: a job that completed hands its search to the idle list
void release_job_device(ps_batch* b, Job* j, bool keep_search = false) {
    if (j->ev) {
        (void)hipEventDestroy(j->ev);
        j->ev = nullptr;
    }
    j->in_flight = false;
    if (j->bfs && keep_search && b->opt.mode == PS_BATCH_STREAMS) {
        IdleSearch x;
        x.puzzle = j->puzzle;
        x.index_count = j->index_count;
        x.footprint = j->footprint;
        x.bfs = j->bfs;
        x.masked_stream = j->masked_stream;
        b->idle.push_back(x);
        j->bfs = nullptr;
        j->masked_stream = nullptr;
    }
    if (j->bfs) {
        ps_bfs_destroy(j->bfs);
        j->bfs = nullptr;
        b->used -= std::min(b->used, j->footprint);
    }
    if (j->masked_stream) {
        (void)hipStreamDestroy(j->masked_stream);
        j->masked_stream = nullptr;
    }
}

void drop_idle(ps_batch* b, size_t i) {
    ps_bfs_destroy(b->idle[i].bfs);
    if (b->idle[i].masked_stream) (void)hipStreamDestroy(b->idle[i].masked_stream);
    b->used -= std::min(b->used, b->idle[i].footprint);
    b->idle.erase(b->idle.begin() + (ptrdiff_t)i);
}

// an idle search for the job, else frees idle searches (oldest first) until
// the job fits the budget; returns whether it fits
bool make_room(ps_batch* b, Job* j, ps_bfs** reuse) {
    *reuse = nullptr;
    for (size_t i = 0; i < b->idle.size(); ++i) {
        if (same_search(b->idle[i].puzzle, b->idle[i].index_count, j->puzzle, j->index_count)) {
            *reuse = b->idle[i].bfs;
            j->masked_stream = b->idle[i].masked_stream;
            b->used -= std::min(b->used, b->idle[i].footprint);
            b->idle.erase(b->idle.begin() + (ptrdiff_t)i);
            return true;
        }
    }
    while (b->used + j->footprint > b->budget && !b->idle.empty()) drop_idle(b, 0);
    return b->used + j->footprint <= b->budget;
}

int admit(ps_batch* b) {
    while (!b->queue.empty() && (int)b->running.size() < b->opt.max_jobs) {
        Job* j = b->queue.front();
        ps_bfs* reuse = nullptr;
        if (!make_room(b, j, &reuse)) break;  // first come, first served
        b->queue.pop_front();
        ps_bfs_options o;
        int r = bfs_opts_for(b, j, &o);
        ps_stream js = b->opt.mode == PS_BATCH_FUSED ? reinterpret_cast<ps_stream>(b->stream) : nullptr;
        if (reuse) {
            j->bfs = reuse;
        } else if (r == PS_OK) {
            if (!b->large_mask.empty() && b->opt.mode == PS_BATCH_STREAMS && (j->states >> LATENCY_BITS) != 0) {
                if (hipExtStreamCreateWithCUMask(&j->masked_stream, (uint32_t)b->large_mask.size(),
                                                 b->large_mask.data()) != hipSuccess) {
                    (void)hipGetLastError();
                    j->masked_stream = nullptr;
                    r = fail(PS_ERR_HIP, "CU-masked stream creation failed");
                }
                js = reinterpret_cast<ps_stream>(j->masked_stream);
            }
            if (r == PS_OK) r = ps_bfs_create(&j->puzzle, &o, js, &j->bfs);
        }
        if (r == PS_OK) {
            b->used += j->footprint;
            r = ps_bfs_reset(j->bfs, j->start);
        }
        if (r == PS_OK && b->opt.mode == PS_BATCH_STREAMS &&
            hipEventCreateWithFlags(&j->ev, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            j->ev = nullptr;
            r = fail(PS_ERR_HIP, "event creation failed");
        }
        if (r != PS_OK) {
            j->error = ps_last_error();
            j->state = PS_JOB_FAILED;
            release_job_device(b, j);
            b->finished_last.push_back(j->id);
            continue;
        }
        j->state = PS_JOB_RUNNING;
        j->t_admit = now_ns();
        j->remaining = j->states;
        b->running.push_back(j);
        b->table_dirty = true;
    }
    return PS_OK;
}

int build_table(ps_batch* b) {
    uint64_t tb = 0, cb = 0;
    uint32_t win = 0;
    for (size_t i = 0; i < b->running.size(); ++i) {
        ps_bfs* x = b->running[i]->bfs;
        psk::MultiJob& J = b->h_jobs[i];
        std::memset(&J, 0, sizeof(J));
        J.e = expand_args(x);
        J.f = finalize_args(x);
        J.l = layer_end_args(x);
        J.kind = (uint32_t)b->running[i]->kind;
        J.slot = (uint32_t)i;
        J.tile_base = tb;
        J.chunk_base = cb;
        tb += x->ntiles;
        cb += x->nchunks;
        if (x->use_window) win = std::max<uint32_t>(win, (uint32_t)(x->tile_words * 2));
    }
    b->total_tiles = tb;
    b->total_chunks = cb;
    b->win_u32_max = win;
    b->lds = ((size_t)win + (MULTI_TPB / 64) * psk::QCAP + psk::MULTI_TL_WORDS) * 4;
    const auto kern = psk::k_multi_expand<MultiKinds, MULTI_TPB>;
    if (b->lds > 65536)
        PS_HIP_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)b->lds));
    int occ = 0;
    PS_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, MULTI_TPB, b->lds));
    if (occ < 1) return fail(PS_ERR_UNSUPPORTED, "multi-search expand does not fit (LDS %zu bytes)", b->lds);
    b->grid_expand = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)occ * b->cus, tb));
    b->grid_fin = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)b->cus * 8, cb));
    PS_HIP_CHECK(hipMemcpyAsync(b->d_jobs, b->h_jobs, b->running.size() * sizeof(psk::MultiJob),
                                hipMemcpyHostToDevice, b->stream));
    b->table_dirty = false;
    return PS_OK;
}

int retire(ps_batch* b, Job* j) {
    ps_bfs* x = j->bfs;
    if (collect_stats(x) != PS_OK) {
        j->error = ps_last_error();
        j->state = PS_JOB_FAILED;
        release_job_device(b, j);
        return PS_OK;
    }
    const int nl = (int)x->h_st->layers_recorded;
    j->counts.resize(nl);
    j->hashes.resize(nl);
    if (nl > 0) {
        // on the job's own stream: a copy on the null stream could wait for other jobs
        PS_HIP_CHECK(hipMemcpyAsync(j->counts.data(), x->layer_counts, (size_t)nl * 8, hipMemcpyDeviceToHost,
                                    x->stream));
        PS_HIP_CHECK(hipMemcpyAsync(j->hashes.data(), x->layer_hashes, (size_t)nl * 8, hipMemcpyDeviceToHost,
                                    x->stream));
        PS_HIP_CHECK(hipStreamSynchronize(x->stream));
    }
    j->expanded = x->stats.expanded;
    j->children = x->stats.children;
    j->state = PS_JOB_DONE;
    j->t_done = now_ns();
    release_job_device(b, j, true);
    b->stats.jobs_done++;
    for (uint64_t c : j->counts) b->stats.states_done += c;
    b->stats.expanded += j->expanded;
    return PS_OK;
}

// Per-job streams without a common wait.  Every running job with nothing in
// flight gets its next layers enqueued on its own stream, followed by a copy of
// its state and its event; layers after the last one are no-ops on the device
// (its kernels see the done flag).  The step then polls the events and returns
// once at least one job has made progress (it is retired if its search ended,
// otherwise its next layers are enqueued at once) or MAX_WAIT_MS have passed,
// so the caller can submit new requests meanwhile.  A job's batch doubles while
// its layers complete within BATCH_FAST_MS and halves when they take longer
// than BATCH_SLOW_MS, which keeps the work queued ahead of a newly admitted job
// on a shared hardware queue short.  A small job thus finishes at its own pace,
// not at the pace of the largest one.
// (batches of 0.3-1 ms rather than 1-3 ms: p50 at 50 jobs/s 4.21 / 3.99 ms
// instead of 4.62 / 4.24 ms on two request streams, the same mean)
constexpr double MAX_WAIT_MS = 2.0, BATCH_FAST_MS = 0.3, BATCH_SLOW_MS = 1.0;

int enqueue_job(Job* j) {
    if (int r = enqueue_layers(j->bfs, j->batch_layers)) return r;
    PS_HIP_CHECK(hipMemcpyAsync(j->bfs->h_st, j->bfs->st, sizeof(BfsDevState), hipMemcpyDeviceToHost,
                                j->bfs->stream));
    PS_HIP_CHECK(hipEventRecord(j->ev, j->bfs->stream));
    j->in_flight = true;
    j->t_enqueue = now_ns();
    return PS_OK;
}

// Shortest first: a job whose remaining states exceed shortest_first times
// those of the smallest running job gets no new layers, for at most
// MAX_HOLD_MS at a time, so a small job does not share the device with a
// large one's persistent kernels while it runs.
constexpr double MAX_HOLD_MS = 20.0;

bool held(const ps_batch* b, Job* j, uint64_t min_remaining, uint64_t now) {
    const int f = b->opt.shortest_first;
    if (f <= 0 || j->remaining <= (uint64_t)f * min_remaining) {
        j->held_since = 0;
        return false;
    }
    if (j->held_since == 0) j->held_since = now;
    if ((double)(now - j->held_since) * 1e-6 < MAX_HOLD_MS) return true;
    j->held_since = 0;  // the hold ran out: one batch of layers, then it may be held again
    return false;
}

int step_streams(ps_batch* b) {
    const auto t0 = std::chrono::steady_clock::now();
    uint64_t min_remaining = ~0ull;
    for (const Job* j : b->running) min_remaining = std::min(min_remaining, j->remaining);
    const uint64_t now0 = now_ns();
    for (Job* j : b->running)
        if (!j->in_flight && !held(b, j, min_remaining, now0))
            if (int r = enqueue_job(j)) return r;
    bool progress = false, retired = false;
    for (int polls = 0; !progress; ++polls) {
        for (Job* j : b->running) {
            if (!j->in_flight) continue;
            const hipError_t q = hipEventQuery(j->ev);
            if (q == hipErrorNotReady) continue;
            PS_HIP_CHECK(q);
            const uint64_t t = now_ns();
            j->in_flight = false;
            j->steps++;
            progress = true;
            const double ms = (double)(t - j->t_enqueue) * 1e-6;
            if (ms < BATCH_FAST_MS) j->batch_layers = std::min(2 * j->batch_layers, MAX_BATCH_LAYERS);
            if (ms > BATCH_SLOW_MS) j->batch_layers = std::max(j->batch_layers / 2, MIN_BATCH_LAYERS);
            const uint64_t visited = j->bfs->h_st->visited;
            j->remaining = j->states > visited ? j->states - visited : 1;
            if (j->bfs->h_st->done) {
                if (int r = retire(b, j)) return r;
                b->finished_last.push_back(j->id);
                retired = true;
            } else if (!held(b, j, min_remaining, t)) {
                if (int r = enqueue_job(j)) return r;
            }
        }
        if (progress ||
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() > MAX_WAIT_MS)
            break;
        if (polls > 64) std::this_thread::yield();
    }
    if (retired) {
        std::vector<Job*> still;
        for (Job* j : b->running)
            if (j->state == PS_JOB_RUNNING) still.push_back(j);
        b->running.swap(still);
    }
    b->stats.steps++;
    b->stats.step_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return PS_OK;
}

int step(ps_batch* b, int32_t* running, int32_t* finished, int32_t* queued) {
    b->finished_last.clear();
    if (int r = admit(b)) return r;
    if (!b->running.empty() && b->opt.mode != PS_BATCH_FUSED) {
        if (int r = step_streams(b)) return r;
    } else if (!b->running.empty()) {
        const auto t0 = std::chrono::steady_clock::now();
        if (b->table_dirty)
            if (int r = build_table(b)) return r;
        psk::MultiArgs m{};
        m.jobs = b->d_jobs;
        m.njobs = (uint32_t)b->running.size();
        m.total_tiles = b->total_tiles;
        m.total_chunks = b->total_chunks;
        m.counters = b->d_counters;
        m.done = b->d_done;
        hipLaunchKernelGGL((psk::k_multi_expand<MultiKinds, MULTI_TPB>), dim3(b->grid_expand), dim3(MULTI_TPB), b->lds,
                           b->stream, m, b->win_u32_max);
        hipLaunchKernelGGL(psk::k_multi_finalize<FIN_TPB>, dim3(b->grid_fin), dim3(FIN_TPB), 0, b->stream, m);
        hipLaunchKernelGGL(psk::k_multi_layer_end<1024>, dim3(m.njobs), dim3(1024), 0, b->stream, m);
        PS_HIP_CHECK(hipGetLastError());
        PS_HIP_CHECK(hipMemcpyAsync(b->h_done, b->d_done, (size_t)m.njobs * 4, hipMemcpyDeviceToHost, b->stream));
        PS_HIP_CHECK(hipStreamSynchronize(b->stream));
        b->stats.steps++;
        b->stats.step_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        std::vector<Job*> still;
        for (size_t i = 0; i < b->running.size(); ++i) {
            Job* j = b->running[i];
            j->steps++;
            if (b->h_done[i]) {
                if (int r = retire(b, j)) return r;
                b->finished_last.push_back(j->id);
                b->table_dirty = true;
            } else {
                still.push_back(j);
            }
        }
        b->running.swap(still);
    }
    if (running) *running = (int32_t)b->running.size();
    if (finished) *finished = (int32_t)b->finished_last.size();
    if (queued) *queued = (int32_t)b->queue.size();
    return PS_OK;
}

int destroy_batch(ps_batch* b) {
    if (!b) return PS_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (auto& kv : b->jobs) {
        release_job_device(b, kv.second);
        delete kv.second;
    }
    while (!b->idle.empty()) drop_idle(b, b->idle.size() - 1);
    if (b->d_jobs) (void)hipFree(b->d_jobs);
    if (b->h_jobs) (void)hipHostFree(b->h_jobs);
    if (b->d_counters) (void)hipFree(b->d_counters);
    if (b->d_done) (void)hipFree(b->d_done);
    if (b->h_done) (void)hipHostFree(b->h_done);
    if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    (void)hipGetLastError();
    return PS_OK;
}

}  // namespace
}  // namespace ps

using namespace ps;

extern "C" {

void ps_batch_options_default(ps_batch_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->max_jobs = 64;
    o->segment_bits = 22;
    o->shortest_first = 4;
}

int ps_batch_create(const ps_batch_options* opt, ps_stream s, ps_batch** out) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (!out) return fail(PS_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    ps_batch_options o;
    if (opt) o = *opt;
    else ps_batch_options_default(&o);
    if (o.max_jobs <= 0) o.max_jobs = 64;
    if (o.max_jobs > 4096) return fail(PS_ERR_INVALID_ARGUMENT, "max_jobs must be at most 4096");
    if (o.segment_bits < 0 || o.segment_bits > 40) return fail(PS_ERR_INVALID_ARGUMENT, "segment_bits out of range");
    if (o.job_grid_blocks < 0) return fail(PS_ERR_INVALID_ARGUMENT, "job_grid_blocks must not be negative");
    if (o.shortest_first < 0) return fail(PS_ERR_INVALID_ARGUMENT, "shortest_first must not be negative");
    if (o.reserved_cus < 0 || o.reserved_cus >= hip_cu_count())
        return fail(PS_ERR_INVALID_ARGUMENT, "reserved_cus must be in [0, %d)", hip_cu_count());
    if (o.mode == PS_BATCH_FUSED)
        return fail(PS_ERR_UNSUPPORTED, "the fused multi-search mode is disabled: its first device run stalled "
                                        "inside a step for a reason not yet found");
    if (o.mode != PS_BATCH_STREAMS) return fail(PS_ERR_INVALID_ARGUMENT, "unknown batch mode %d", o.mode);
    ps_batch* b = new (std::nothrow) ps_batch();
    if (!b) return fail(PS_ERR_OUT_OF_MEMORY, "host allocation failed");
    auto cleanup = [&](int rc) {
        destroy_batch(b);
        return rc;
    };
    b->opt = o;
    b->cus = (uint32_t)hip_cu_count();
    if (s) {
        b->stream = stream_of(s);
    } else {
        if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
            return cleanup(fail(PS_ERR_HIP, "stream creation failed"));
        b->own_stream = true;
    }
    if (o.hbm_budget) {
        b->budget = o.hbm_budget;
    } else {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return cleanup(fail(PS_ERR_HIP, "hipMemGetInfo failed"));
        b->budget = free_b / 2;
    }
    const size_t jb = (size_t)o.max_jobs * sizeof(psk::MultiJob);
    if (hipMalloc(&b->d_jobs, jb) != hipSuccess || hipMalloc(&b->d_counters, 64) != hipSuccess ||
        hipMalloc(&b->d_done, (size_t)o.max_jobs * 4) != hipSuccess) {
        (void)hipGetLastError();
        return cleanup(fail(PS_ERR_OUT_OF_MEMORY, "device allocation failed"));
    }
    if (hipHostMalloc((void**)&b->h_jobs, jb, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&b->h_done, (size_t)o.max_jobs * 4, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return cleanup(fail(PS_ERR_OUT_OF_MEMORY, "pinned host allocation failed"));
    }
    if (o.reserved_cus > 0) {
        b->large_mask.assign((b->cus + 31) / 32, 0u);
        for (uint32_t cu = (uint32_t)o.reserved_cus; cu < b->cus; ++cu) b->large_mask[cu / 32] |= 1u << (cu % 32);
    }
    if (int r = launch_fill(b->d_counters, 0, 64, b->stream)) return cleanup(r);
    if (hipStreamSynchronize(b->stream) != hipSuccess) return cleanup(fail(PS_ERR_HIP, "stream synchronize failed"));
    *out = b;
    return PS_OK;
    PS_API_END
}

int ps_batch_destroy(ps_batch* b) {
    PS_API_BEGIN
    return destroy_batch(b);
    PS_API_END
}

int ps_batch_submit(ps_batch* b, const ps_puzzle* p, uint64_t start, uint64_t index_count, uint64_t* job_id) {
    PS_API_BEGIN
    if (!b || !p) return fail(PS_ERR_INVALID_ARGUMENT, "NULL argument");
    const int kind = multi_kind_of(*p);
    if (kind < 0)
        return fail(PS_ERR_UNSUPPORTED, "puzzle kind=%d size=%d height=%d is not in the multi-search table", p->kind,
                    p->size, p->height);
    ps_puzzle_info info{};
    if (int r = puzzle_query(p, &info)) return r;
    const uint64_t n = index_count ? index_count : info.indexes_count;
    if (start >= n)
        return fail(PS_ERR_INDEX_RANGE, "start index %llu outside [0, %llu)", (unsigned long long)start,
                    (unsigned long long)n);
    Job* j = new (std::nothrow) Job();
    if (!j) return fail(PS_ERR_OUT_OF_MEMORY, "host allocation failed");
    j->id = b->next_id++;
    j->puzzle = *p;
    j->start = start;
    j->index_count = index_count;
    j->kind = kind;
    ps_bfs_options o;
    if (int r = bfs_opts_for(b, j, &o)) {
        delete j;
        return r;
    }
    j->states = n;
    if (int r = ps_bfs_footprint(p, &o, &j->footprint)) {
        delete j;
        return r;
    }
    if (j->footprint > b->budget) {
        delete j;
        return fail(PS_ERR_CAPACITY, "job needs %llu bytes of HBM, the budget is %llu",
                    (unsigned long long)j->footprint, (unsigned long long)b->budget);
    }
    j->t_submit = now_ns();
    b->jobs[j->id] = j;
    b->queue.push_back(j);
    if (job_id) *job_id = j->id;
    return PS_OK;
    PS_API_END
}

int ps_batch_step(ps_batch* b, int32_t* running, int32_t* finished, int32_t* queued) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (!b) return fail(PS_ERR_INVALID_ARGUMENT, "NULL batch");
    return step(b, running, finished, queued);
    PS_API_END
}

int ps_batch_finished(const ps_batch* b, uint64_t* ids, int32_t cap, int32_t* n) {
    PS_API_BEGIN
    if (!b || !n) return fail(PS_ERR_INVALID_ARGUMENT, "NULL argument");
    *n = (int32_t)b->finished_last.size();
    if (ids)
        for (int32_t i = 0; i < std::min<int32_t>(cap, *n); ++i) ids[i] = b->finished_last[i];
    return PS_OK;
    PS_API_END
}

int ps_batch_job(const ps_batch* b, uint64_t job_id, ps_batch_job_info* info, uint64_t* counts, uint64_t* hashes,
                 int32_t cap) {
    PS_API_BEGIN
    if (!b || !info) return fail(PS_ERR_INVALID_ARGUMENT, "NULL argument");
    auto it = b->jobs.find(job_id);
    if (it == b->jobs.end()) return fail(PS_ERR_INVALID_ARGUMENT, "unknown job %llu", (unsigned long long)job_id);
    const Job* j = it->second;
    std::memset(info, 0, sizeof(*info));
    info->state = j->state;
    info->layers = (int32_t)j->counts.size();
    info->footprint = j->footprint;
    info->steps = j->steps;
    info->expanded = j->expanded;
    info->children = j->children;
    info->submit_ns = j->t_submit;
    info->admit_ns = j->t_admit;
    info->done_ns = j->t_done;
    const int32_t nc = std::min<int32_t>(cap, (int32_t)j->counts.size());
    if (counts && nc > 0) std::memcpy(counts, j->counts.data(), (size_t)nc * 8);
    if (hashes && nc > 0) std::memcpy(hashes, j->hashes.data(), (size_t)nc * 8);
    if (j->state == PS_JOB_FAILED) return fail(PS_ERR_HIP, "job %llu failed: %s", (unsigned long long)job_id,
                                               j->error.c_str());
    return PS_OK;
    PS_API_END
}

int ps_batch_forget(ps_batch* b, uint64_t job_id) {
    PS_API_BEGIN
    if (!b) return fail(PS_ERR_INVALID_ARGUMENT, "NULL batch");
    auto it = b->jobs.find(job_id);
    if (it == b->jobs.end()) return PS_OK;
    Job* j = it->second;
    if (j->state == PS_JOB_QUEUED || j->state == PS_JOB_RUNNING)
        return fail(PS_ERR_INVALID_ARGUMENT, "job %llu has not finished", (unsigned long long)job_id);
    b->jobs.erase(it);
    delete j;
    return PS_OK;
    PS_API_END
}

int ps_batch_stats_get(const ps_batch* b, ps_batch_stats* out) {
    PS_API_BEGIN
    if (!b || !out) return fail(PS_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = b->stats;
    out->hbm_used = b->used;
    out->hbm_budget = b->budget;
    return PS_OK;
    PS_API_END
}

}  // extern "C"
