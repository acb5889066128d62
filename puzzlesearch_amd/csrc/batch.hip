// Multi-search engine: continuous batching of complete three-bit BFS jobs.
//
// Jobs of different puzzles and sizes are submitted at any time.  A step
//   1 admits queued jobs, first come first served, while their HBM footprint
//     (ps_bfs_footprint) fits the budget and a job slot is free;
//   2 advances every running job by one layer with three launches shared by all
//     of them: a persistent expand over the concatenated tile lists, a
//     persistent finalize over the concatenated chunk lists and a bookkeeping
//     kernel with one block per job (k_multi_* in bfs_kernels.h);
//   3 retires the jobs whose search ended: their layer counts and frontier
//     hashes are copied out and their device memory is released at once,
//     so queued jobs can be admitted in the next step.
// Two ways of running a step:
//   PS_BATCH_STREAMS (default): every job is a ps_bfs on its own stream; one
//     layer of every running job is enqueued before any of them is waited for,
//     so the jobs' kernels overlap on the device;
//   PS_BATCH_FUSED: the three multi-search launches above.  Disabled: its first
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
    int kind = -1;
    int state = PS_JOB_QUEUED;
    ps_bfs* bfs = nullptr;
    uint64_t t_submit = 0, t_admit = 0, t_done = 0;
    uint64_t steps = 0;
    std::vector<uint64_t> counts, hashes;
    uint64_t expanded = 0, children = 0;
    std::string error;
};

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
        o->cross_marks = PS_CROSS_ATOMIC;
        if (j->puzzle.kind == PS_PANCAKE && j->puzzle.size >= 10 && j->puzzle.size <= 12 &&
            n >= factorial(j->puzzle.size))
            o->cross_marks = PS_CROSS_LISTS;
    }
    // segments pay off once a search is large; small ones run as one segment
    o->segment_bits = (b->opt.segment_bits > 0 && n > (1ull << b->opt.segment_bits)) ? b->opt.segment_bits : 0;
    return PS_OK;
}

void release_job_device(ps_batch* b, Job* j) {
    if (j->bfs) {
        ps_bfs_destroy(j->bfs);
        j->bfs = nullptr;
        b->used -= std::min(b->used, j->footprint);
    }
}

int admit(ps_batch* b) {
    while (!b->queue.empty() && (int)b->running.size() < b->opt.max_jobs) {
        Job* j = b->queue.front();
        if (b->used + j->footprint > b->budget) break;  // first come, first served
        b->queue.pop_front();
        ps_bfs_options o;
        int r = bfs_opts_for(b, j, &o);
        ps_stream js = b->opt.mode == PS_BATCH_FUSED ? reinterpret_cast<ps_stream>(b->stream) : nullptr;
        if (r == PS_OK) r = ps_bfs_create(&j->puzzle, &o, js, &j->bfs);
        if (r == PS_OK) {
            b->used += j->footprint;
            r = ps_bfs_reset(j->bfs, j->start);
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
        PS_HIP_CHECK(hipMemcpy(j->counts.data(), x->layer_counts, (size_t)nl * 8, hipMemcpyDeviceToHost));
        PS_HIP_CHECK(hipMemcpy(j->hashes.data(), x->layer_hashes, (size_t)nl * 8, hipMemcpyDeviceToHost));
    }
    j->expanded = x->stats.expanded;
    j->children = x->stats.children;
    j->state = PS_JOB_DONE;
    j->t_done = now_ns();
    release_job_device(b, j);
    b->stats.jobs_done++;
    for (uint64_t c : j->counts) b->stats.states_done += c;
    b->stats.expanded += j->expanded;
    return PS_OK;
}

// one layer of every running job, each on its own stream, then the done flags
int step_streams(ps_batch* b) {
    const auto t0 = std::chrono::steady_clock::now();
    for (Job* j : b->running) {
        if (int r = launch_batched_layer(j->bfs, false)) return r;
        PS_HIP_CHECK(hipMemcpyAsync(j->bfs->h_st, j->bfs->st, sizeof(BfsDevState), hipMemcpyDeviceToHost,
                                    j->bfs->stream));
    }
    std::vector<Job*> still;
    for (Job* j : b->running) {
        PS_HIP_CHECK(hipStreamSynchronize(j->bfs->stream));
        j->steps++;
        if (j->bfs->h_st->done) {
            if (int r = retire(b, j)) return r;
            b->finished_last.push_back(j->id);
        } else {
            still.push_back(j);
        }
    }
    b->running.swap(still);
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
