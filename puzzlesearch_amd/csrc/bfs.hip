// Device-resident BFS driver: InMemoryThreeBitBFS / InMemoryTwoBitBFS
// (Common/InMemoryThreeBitBFS.cpp:4-48, Common/InMemoryTwoBitBFS.cpp:24-69)
// with the whole layer loop on the GPU, optionally split into segments of
// 2^segment_bits states as DiskBasedOptThreeBitBFS does
// (Common/DiskBasedOptThreeBitBFS.cpp:54-144).
//
// Two execution modes:
//  * PS_BFS_BATCHED: per layer one persistent expand kernel over every tile of
//    every live segment, one persistent finalize kernel, one bookkeeping kernel;
//    the layer loop is captured in a hipGraph (graph_layers layers per replay)
//    and the host only polls the done flag between replays.
//  * PS_BFS_PLAIN: the host walks the segments like the reference's segment
//    loop, launching one expand and one finalize kernel per live segment per
//    layer.
//
// Algorithmic bytes (for the roofline) per layer: 8 bytes per cur word scanned,
// 8 bytes per child generated (one 32-bit read-modify-write), and for each
// finalized word 8*(4 or 5) bytes (read new, old [, cur], write new, old), or 8
// bytes for a word of old that is only cleared.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <new>
#include <stdexcept>
#include <vector>

#include "bfs_internal.h"
#include "bfs_launch.h"
#include "ps_internal.h"


namespace {

uint64_t host_splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

uint64_t host_word_hash(uint64_t j, uint64_t w) { return w ? host_splitmix64(host_splitmix64(j) ^ w) : 0; }

constexpr int MAX_LAYERS_CAP = ps::BFS_MAX_LAYERS;

// The start of a search: the start state's word of the first array, its
// segment's count, layer 0's count and hash, the first tile list entry and the
// device state, written in order on the search's own stream (kernel arguments,
// so no host buffer has to outlive the launch).
struct ResetStart {
    uint64_t* word_ptr;
    uint64_t word;
    unsigned long long* seg_count;  // NULL: none
    unsigned long long* layer_counts;
    unsigned long long* layer_hashes;
    uint64_t* visited_word;  // NULL: no visited array
    unsigned long long h0;
    uint32_t* tile_list;  // NULL: none
    uint32_t tile;
    psk::BfsDevState* st;
    psk::BfsDevState st_value;
};

__global__ void k_bfs_reset_start(ResetStart r) {
    if (threadIdx.x != 0) return;
    *r.word_ptr = r.word;
    if (r.visited_word) *r.visited_word = r.word;
    if (r.seg_count) *r.seg_count = 1ull;
    *r.layer_counts = 1ull;
    *r.layer_hashes = r.h0;
    if (r.tile_list) *r.tile_list = r.tile;
    *r.st = r.st_value;
}

}  // namespace


namespace ps {

psk::ExpandArgs expand_args(ps_bfs* b) {
    psk::ExpandArgs a{};
    a.prm = params_of(b->puzzle);
    for (int i = 0; i < 3; ++i) a.bits[i] = b->arr[i];
    a.nbits = b->nbits;
    a.nwords = b->nwords;
    a.tile_words = b->tile_words;
    a.tile_begin = 0;
    a.tile_end = b->ntiles;
    a.word_lo = 0;
    a.word_hi = b->nwords;
    a.seg_shift = b->seg_shift;
    a.nseg = b->nseg;
    a.seg_count = b->seg_count;
    a.touched = b->touched;
    a.st = b->st;
    a.persistent = 1;
    a.use_window = b->use_window;
    a.pull_capable = b->pull_capable;
    a.xmark = b->xmark;
    a.skip_visited = b->opt.skip_visited;
    return a;
}

psk::FinalizeArgs finalize_args(ps_bfs* b) {
    psk::FinalizeArgs f{};
    for (int i = 0; i < 3; ++i) f.bits[i] = b->arr[i];
    f.nwords = b->nwords;
    f.seg_words = b->seg_words;
    f.chunks_per_seg = b->chunks_per_seg;
    f.chunk_begin = 0;
    f.chunk_end = b->nchunks;
    f.nseg = b->nseg;
    f.odd = b->info.has_odd_cycles;
    f.hash = b->opt.hash;
    f.persistent = 1;
    f.seg_count = b->seg_count;
    f.touched = b->touched;
    f.st = b->st;
    f.xmark = b->xmark;
    return f;
}

psk::ListArgs list_args(ps_bfs* b) {
    psk::ListArgs x{};
    x.lists = b->xlists;
    x.counts = b->xcounts;
    x.flags = b->xflags;
    x.tiles = b->xtiles;
    return x;
}

psk::ApplyArgs apply_args(ps_bfs* b) {
    psk::ApplyArgs a{};
    a.x = list_args(b);
    for (int i = 0; i < 3; ++i) a.bits[i] = b->arr[i];
    a.seg_shift = b->seg_shift;
    a.nseg = b->nseg;
    a.touched = b->touched;
    a.st = b->st;
    a.persistent = 1;
    a.tile_begin = 0;
    a.tile_end = b->xtiles;
    return a;
}

psk::StepArgs step_args(ps_bfs* b, bool first) {
    psk::StepArgs a{};
    for (int i = 0; i < 3; ++i) a.bits[i] = b->arr[i];
    a.st = b->st;
    a.lists[0] = b->xlists;
    a.lists[1] = b->xlists1;
    a.counts[0] = b->xcounts;
    a.counts[1] = b->xcounts1;
    a.xflags[0] = b->xflags;
    a.xflags[1] = b->xflags1;
    a.wflags[0] = b->wflags;
    a.wflags[1] = b->wflags + b->xtiles;
    a.tcount = b->tcount;
    a.active = b->active;
    a.tiles = b->xtiles;
    // a start outside the tiles (an index space past n!) leaves nothing to expand
    const uint64_t t0 = b->start_index / factorial(9);
    a.first_tile = (uint32_t)std::min<uint64_t>(t0, b->xtiles);
    a.odd = b->info.has_odd_cycles;
    a.hash = b->opt.hash;
    a.first = first ? 1 : 0;
    a.role = 0;
    a.visited = b->visited;
    a.nbits = b->nbits;
    return a;
}

psk::PullArgs pull_args(ps_bfs* b) {
    psk::PullArgs a{};
    for (int i = 0; i < 3; ++i) a.bits[i] = b->arr[i];
    a.st = b->st;
    a.visited = b->visited;
    a.wflags[0] = b->wflags;
    a.wflags[1] = b->wflags + b->xtiles;
    a.tiles = b->xtiles;
    a.nbits = b->nbits;
    return a;
}

psk::LayerEndArgs layer_end_args(ps_bfs* b) {
    psk::LayerEndArgs e{};
    e.nbits = b->nbits;
    e.nseg = b->nseg;
    e.seg_count = b->seg_count;
    e.touched = b->touched;
    e.layer_counts = b->layer_counts;
    e.layer_hashes = b->layer_hashes;
    e.st = b->st;
    e.bits = b->bits;
    e.active = b->tile_list;
    e.ntiles = b->ntiles;
    e.tile_bits = b->tile_words * 64;
    e.seg_shift = b->seg_shift;
    e.rot4 = b->tstep;
    e.list_only = 0;
    e.touched_par[0] = b->touched;
    e.touched_par[1] = b->touched1;
    return e;
}

psk::TileStepArgs tile_step_args(ps_bfs* b) {
    psk::TileStepArgs t{};
    for (int i = 0; i < 4; ++i) t.bits[i] = b->arr[i];
    t.seg_count = b->seg_count;
    t.touched[0] = b->touched;
    t.touched[1] = b->touched1;
    t.seg_words = b->seg_words;
    t.odd = b->info.has_odd_cycles;
    t.hash = b->opt.hash;
    return t;
}

}  // namespace ps

namespace ps {
namespace {
int record(ps_bfs* b, size_t* idx);
}  // namespace

int launch_layer_end(ps_bfs* b) {
    if (b->fused) return pkl_layer_end(layer_end_args(b), step_args(b, false), 1, b->stream);
    hipLaunchKernelGGL(psk::k_bfs_layer_end<1024>, dim3(1), dim3(1024), 0, b->stream, layer_end_args(b));
    PS_HIP_CHECK(hipGetLastError());
    return PS_OK;
}

// One batched layer: persistent expand, persistent finalize, bookkeeping.
int launch_batched_layer(ps_bfs* b, bool timed) {
    if (b->bits == 2) {
        psk::Bfs2Args a2{};
        a2.prm = params_of(b->puzzle);
        a2.arr = b->arr[0];
        a2.nbits = b->nbits;
        a2.nwords = b->nwords;
        a2.tile_words = b->tile_words;
        a2.ntiles = b->ntiles;
        a2.st = b->st;
        size_t e0 = 0, e1 = 0;
        if (timed)
            if (int r = record(b, &e0)) return r;
        LaunchCfg c{256, b->grid_bfs2, (size_t)4 * psk::QCAP * 4, b->stream};
        if (int r = bfs2_layer(b->puzzle, a2, c, nullptr)) return r;
        if (timed) {
            if (int r = record(b, &e1)) return r;
            if (b->collect_kernel_times) b->expand_spans.push_back({e0, e1});
        }
        b->stats.expand_launches++;
        if (b->opt.hash) {
            uint32_t g = (uint32_t)std::min<uint64_t>((b->nwords + 1023) / 1024, 4096);
            if (g == 0) g = 1;
            hipLaunchKernelGGL(psk::k_bfs2_hash<256>, dim3(g), dim3(256), 0, b->stream, (const uint64_t*)b->arr[0],
                               b->nwords, b->st);
            PS_HIP_CHECK(hipGetLastError());
        }
        return launch_layer_end(b);
    }
    if (b->tstep) {
        if (!b->primed)
            if (int r = prime_tstep(b)) return r;
        size_t t0 = 0, t1 = 0;
        if (timed)
            if (int r = record(b, &t0)) return r;
        psk::ExpandArgs a = expand_args(b);
        a.active = b->tile_list;
        LaunchCfg ct{b->tpb_expand, b->grid_tstep, b->lds_expand, b->stream};
        if (int r = bfs3_tile_step(b->puzzle, a, tile_step_args(b), ct, nullptr)) return r;
        if (timed) {
            if (int r = record(b, &t1)) return r;
            if (b->collect_kernel_times) b->expand_spans.push_back({t0, t1});
        }
        b->stats.expand_launches++;
        return launch_layer_end(b);
    }
    if (b->fused) {
        if (!b->primed)
            if (int r = prime_fused(b)) return r;
        size_t f0 = 0, fm = 0, f1 = 0;
        if (timed)
            if (int r = record(b, &f0)) return r;
        LaunchCfg cs{b->tpb_expand, b->grid_step, b->lds_step, b->stream};
        if (int r = pkl_step(b->puzzle, step_args(b, false), cs, nullptr)) return r;
        b->stats.expand_launches++;
        if (b->visited) {
            // a split layer: its expansion top-down (role 1) or bottom-up; each
            // of the two exits at once unless the device chose it
            psk::StepArgs a1 = step_args(b, false);
            a1.role = 1;
            if (int r = pkl_step(b->puzzle, a1, cs, nullptr)) return r;
            b->stats.expand_launches++;
            if (timed)
                if (int r = record(b, &fm)) return r;
            LaunchCfg cp{b->tpb_expand, b->grid_pull, b->lds_pull, b->stream};
            if (int r = pkl_pull(b->puzzle, pull_args(b), cp, nullptr)) return r;
            b->stats.pull_launches++;
        }
        if (timed) {
            if (int r = record(b, &f1)) return r;
            if (b->collect_kernel_times) {
                // k_pkl_step launches and the k_pkl_pull launch timed apart
                if (b->visited) {
                    b->expand_spans.push_back({f0, fm});
                    b->pull_spans.push_back({fm, f1});
                } else {
                    b->expand_spans.push_back({f0, f1});
                }
            }
        }
        return launch_layer_end(b);
    }
    size_t e0 = 0, e1 = 0, e2 = 0, ea = 0;
    if (timed)
        if (int r = record(b, &e0)) return r;
    LaunchCfg ce{b->tpb_expand, b->grid_expand, b->lds_expand, b->stream};
    if (b->lists) {
        psk::ExpandArgs a = expand_args(b);
        a.tile_end = std::min<uint64_t>(a.tile_end, b->xtiles);
        if (int r = pkl_expand(b->puzzle, a, list_args(b), ce, nullptr)) return r;
    } else {
        psk::ExpandArgs a = expand_args(b);
        a.active = b->tile_list;
        if (int r = bfs3_expand(b->puzzle, a, ce, nullptr)) return r;
    }
    if (timed)
        if (int r = record(b, &e1)) return r;
    size_t efin = e1;  // start of the finalize span
    if (b->lists) {
        LaunchCfg ca{b->tpb_apply, b->grid_apply, b->lds_apply, b->stream};
        if (int r = pkl_apply(b->puzzle, apply_args(b), ca, nullptr)) return r;
        b->stats.apply_launches++;
        if (timed) {
            if (int r = record(b, &ea)) return r;
            if (b->collect_kernel_times) b->apply_spans.push_back({e1, ea});
            efin = ea;
        }
    }
    hipLaunchKernelGGL(psk::k_bfs3_finalize<256>, dim3(b->grid_fin), dim3(256), 0, b->stream, finalize_args(b));
    PS_HIP_CHECK(hipGetLastError());
    if (timed) {
        if (int r = record(b, &e2)) return r;
        if (b->collect_kernel_times) {
            b->expand_spans.push_back({e0, e1});
            b->fin_spans.push_back({efin, e2});
        }
    }
    b->stats.expand_launches++;
    b->stats.finalize_launches++;
    return launch_layer_end(b);
}


namespace {

int take_event(ps_bfs* b, hipEvent_t* out, size_t* idx) {
    if (b->ev_used == b->ev_pool.size()) {
        hipEvent_t e;
        PS_HIP_CHECK(hipEventCreate(&e));
        b->ev_pool.push_back(e);
    }
    *idx = b->ev_used;
    *out = b->ev_pool[b->ev_used++];
    return PS_OK;
}

int record(ps_bfs* b, size_t* idx) {
    if (!b->collect_kernel_times) return PS_OK;
    hipEvent_t e;
    if (int r = take_event(b, &e, idx)) return r;
    PS_HIP_CHECK(hipEventRecord(e, b->stream));
    return PS_OK;
}

}  // namespace

int prime_tstep(ps_bfs* b) {
    size_t f0 = 0, f1 = 0;
    if (int r = record(b, &f0)) return r;
    // the start layer, expanded from arr[0] into arr[1] (the three-array roles
    // at layer 0), its marks flagged in the touched bits of odd mark layers
    psk::ExpandArgs a = expand_args(b);
    a.active = b->tile_list;
    a.touched = b->touched1;
    LaunchCfg ce{b->tpb_expand, b->grid_expand, b->lds_expand, b->stream};
    if (int r = bfs3_expand(b->puzzle, a, ce, nullptr)) return r;
    if (int r = record(b, &f1)) return r;
    if (b->collect_kernel_times) b->expand_spans.push_back({f0, f1});
    b->stats.expand_launches++;
    psk::LayerEndArgs e = layer_end_args(b);
    e.list_only = 1;
    hipLaunchKernelGGL(psk::k_bfs_layer_end<1024>, dim3(1), dim3(1024), 0, b->stream, e);
    PS_HIP_CHECK(hipGetLastError());
    b->primed = true;
    return PS_OK;
}

int prime_fused(ps_bfs* b) {
    size_t f0 = 0, f1 = 0;
    if (int r = record(b, &f0)) return r;
    LaunchCfg cs{b->tpb_expand, b->grid_step, b->lds_step, b->stream};
    if (int r = pkl_step(b->puzzle, step_args(b, true), cs, nullptr)) return r;
    if (int r = record(b, &f1)) return r;
    if (b->collect_kernel_times) b->expand_spans.push_back({f0, f1});
    b->stats.expand_launches++;
    if (int r = pkl_layer_end(layer_end_args(b), step_args(b, false), 0, b->stream)) return r;
    b->primed = true;
    return PS_OK;
}

namespace {

// One plain layer: host reads the per-segment counts and launches one expand
// and one finalize kernel per segment that needs it.
int run_plain_layer(ps_bfs* b, bool* done) {
    const uint32_t L = b->h_st->layer;
    (void)L;
    PS_HIP_CHECK(hipMemcpyAsync(b->h_st, b->st, sizeof(BfsDevState), hipMemcpyDeviceToHost, b->stream));
    PS_HIP_CHECK(hipMemcpyAsync(b->h_seg, b->seg_count, (size_t)3 * b->nseg * 8, hipMemcpyDeviceToHost, b->stream));
    PS_HIP_CHECK(hipStreamSynchronize(b->stream));
    if (b->h_st->done) {
        *done = true;
        return PS_OK;
    }
    const uint32_t layer = b->h_st->layer;
    const unsigned long long* sc_cur = b->h_seg + (uint64_t)(layer % 3) * b->nseg;
    const unsigned long long* sc_old = b->h_seg + (uint64_t)((layer + 2) % 3) * b->nseg;
    size_t e0 = 0, e1 = 0, e2 = 0;
    if (int r = record(b, &e0)) return r;
    psk::ExpandArgs a = expand_args(b);
    a.persistent = 0;
    for (uint32_t s = 0; s < b->nseg; ++s) {
        if (b->nseg > 1 && sc_cur[s] == 0) continue;
        const uint64_t w0 = (uint64_t)s * b->seg_words;
        const uint64_t w1 = std::min(w0 + b->seg_words, b->nwords);
        a.word_lo = w0;
        a.word_hi = w1;
        a.tile_begin = w0 / b->tile_words;
        a.tile_end = (w1 + b->tile_words - 1) / b->tile_words;
        if (b->lists) a.tile_end = std::min<uint64_t>(a.tile_end, b->xtiles);
        if (a.tile_end <= a.tile_begin) continue;
        LaunchCfg c{b->tpb_expand, (uint32_t)(a.tile_end - a.tile_begin), b->lds_expand, b->stream};
        if (b->lists) {
            if (int r = pkl_expand(b->puzzle, a, list_args(b), c, nullptr)) return r;
        } else {
            if (int r = bfs3_expand(b->puzzle, a, c, nullptr)) return r;
        }
        b->stats.expand_launches++;
    }
    if (int r = record(b, &e1)) return r;
    size_t efin = e1;  // start of the finalize span
    if (b->lists) {
        // every list of this layer is complete: one pass over all tiles
        LaunchCfg ca{b->tpb_apply, b->grid_apply, b->lds_apply, b->stream};
        if (int r = pkl_apply(b->puzzle, apply_args(b), ca, nullptr)) return r;
        b->stats.apply_launches++;
        if (int r = record(b, &efin)) return r;
        if (b->collect_kernel_times) b->apply_spans.push_back({e1, efin});
    }
    if (b->nseg > 1) {
        PS_HIP_CHECK(hipMemcpyAsync(b->h_touched, b->touched, (size_t)((b->nseg + 31) / 32) * 4,
                                    hipMemcpyDeviceToHost, b->stream));
        PS_HIP_CHECK(hipStreamSynchronize(b->stream));
    }
    psk::FinalizeArgs f = finalize_args(b);
    f.persistent = 0;
    for (uint32_t s = 0; s < b->nseg; ++s) {
        if (b->nseg > 1) {
            const bool touched = (b->h_touched[s >> 5] >> (s & 31)) & 1u;
            if (!touched && sc_old[s] == 0) continue;
        }
        f.chunk_begin = (uint64_t)s * b->chunks_per_seg;
        f.chunk_end = f.chunk_begin + b->chunks_per_seg;
        hipLaunchKernelGGL(psk::k_bfs3_finalize<256>, dim3((uint32_t)b->chunks_per_seg), dim3(256), 0, b->stream, f);
        PS_HIP_CHECK(hipGetLastError());
        b->stats.finalize_launches++;
    }
    if (int r = record(b, &e2)) return r;
    if (b->collect_kernel_times) {
        b->expand_spans.push_back({e0, e1});
        b->fin_spans.push_back({e1, e2});
    }
    return launch_layer_end(b);
}

int build_graph(ps_bfs* b) {
    if (b->graph_exec) return PS_OK;
    hipGraph_t graph = nullptr;
    PS_HIP_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = PS_OK;
    for (int i = 0; i < b->graph_layers && rc == PS_OK; ++i) rc = launch_batched_layer(b, false);
    hipError_t e = hipStreamEndCapture(b->stream, &graph);
    if (rc != PS_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    PS_HIP_CHECK(e);
    hipError_t ie = hipGraphInstantiate(&b->graph_exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    PS_HIP_CHECK(ie);
    // the launches above were only captured; do not count them
    b->stats.expand_launches = 0;
    b->stats.finalize_launches = 0;
    b->stats.apply_launches = 0;
    return PS_OK;
}

int destroy_bfs(ps_bfs* b) {
    if (!b) return PS_OK;
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    for (auto e : b->ev_pool) (void)hipEventDestroy(e);
    if (b->ev_run0) (void)hipEventDestroy(b->ev_run0);
    if (b->ev_run1) (void)hipEventDestroy(b->ev_run1);
    // the two-bit search aliases arr[1] and arr[2] to arr[0]
    if (b->arr[3]) (void)hipFree(b->arr[3]);
    if (b->touched1) (void)hipFree(b->touched1);
    for (int i = 0; i < 3; ++i) {
        bool alias = false;
        for (int j = 0; j < i; ++j) alias = alias || b->arr[j] == b->arr[i];
        if (b->arr[i] && !alias) (void)hipFree(b->arr[i]);
    }
    if (b->xmark) (void)hipFree(b->xmark);
    if (b->tile_list) (void)hipFree(b->tile_list);
    if (b->xlists) (void)hipFree(b->xlists);
    if (b->xcounts) (void)hipFree(b->xcounts);
    if (b->xflags) (void)hipFree(b->xflags);
    if (b->xlists1) (void)hipFree(b->xlists1);
    if (b->visited) (void)hipFree(b->visited);
    if (b->xcounts1) (void)hipFree(b->xcounts1);
    if (b->xflags1) (void)hipFree(b->xflags1);
    if (b->wflags) (void)hipFree(b->wflags);
    if (b->tcount) (void)hipFree(b->tcount);
    if (b->active) (void)hipFree(b->active);
    if (b->seg_count) (void)hipFree(b->seg_count);
    if (b->touched) (void)hipFree(b->touched);
    if (b->layer_counts) (void)hipFree(b->layer_counts);
    if (b->layer_hashes) (void)hipFree(b->layer_hashes);
    if (b->st) (void)hipFree(b->st);
    if (b->h_st) (void)hipHostFree(b->h_st);
    if (b->h_seg) (void)hipHostFree(b->h_seg);
    if (b->h_touched) (void)hipHostFree(b->h_touched);
    if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    // a failed release above must not surface as the error of a later, unrelated call
    (void)hipGetLastError();
    return PS_OK;
}

// A bottom-up layer tests every unvisited index for a neighbour in the
// frontier, so it is exact only where every index of the space is a state of
// the searched graph.  That excludes symmetry-reduced Hanoi (non-canonical
// indexes), the packed sliding index (unused (segment, blank) codes) and the
// pancake index of more than 12 pancakes (29-bit segments of 12! states).  An
// index space of nbits indexes must also end at the last state: the 12-pancake
// IndexesCount() is 2^29 > 12!, and an index >= n! would be ranked modulo n!
// and inherit the neighbours of a real state.
bool every_index_a_state(const ps_puzzle& p, uint64_t nbits) {
    switch (p.kind) {
    case PS_PANCAKE: return p.size <= 12 && nbits <= factorial(p.size);
    case PS_SLIDING: return nbits <= factorial(p.size * p.height);
    case PS_HANOI4:
    case PS_HANOI4_OPT: return (p.flags & PS_FLAG_SYMMETRY) == 0 && nbits <= (1ull << (2 * p.size));
    }
    return false;
}

struct Geometry {
    uint64_t nbits, nwords, seg_words;
    uint32_t nseg;
    int seg_shift;
};

int geometry(const ps_puzzle* p, const ps_bfs_options* o, const ps_puzzle_info& info, Geometry* g) {
    g->nbits = o->index_count ? o->index_count : info.indexes_count;
    if (g->nbits == 0) return fail(PS_ERR_INVALID_ARGUMENT, "empty index space");
    if (o->index_count && o->index_count < info.indexes_count && p->kind != PS_PANCAKE && p->kind != PS_SLIDING)
        return fail(PS_ERR_INVALID_ARGUMENT, "index_count smaller than the puzzle's index space");
    if (o->bits == 2) {
        g->nwords = (g->nbits + 31) / 32;
        g->nseg = 1;
        g->seg_words = g->nwords;
        g->seg_shift = 63;
        if (o->segment_bits) return fail(PS_ERR_UNSUPPORTED, "the two-bit search runs as one segment");
        return PS_OK;
    }
    g->nwords = (g->nbits + 63) / 64;
    if (o->segment_bits == 0) {
        g->nseg = 1;
        g->seg_words = g->nwords;
        g->seg_shift = 63;
    } else {
        if (o->segment_bits < 12 || o->segment_bits > 40)
            return fail(PS_ERR_INVALID_ARGUMENT, "segment_bits %d not in [12, 40]", o->segment_bits);
        g->seg_shift = o->segment_bits;
        g->seg_words = 1ull << (o->segment_bits - 6);
        const uint64_t ns = (g->nwords + g->seg_words - 1) / g->seg_words;
        if (ns > (1ull << 22)) return fail(PS_ERR_INVALID_ARGUMENT, "too many segments (%llu)", (unsigned long long)ns);
        g->nseg = (uint32_t)ns;
    }
    return PS_OK;
}

// Cross-tile lists of n pancakes: n!/9! tiles, each (n-9)*9! entries and
// one fill count per list.
void list_geometry(int n, uint32_t* tiles, uint32_t* slots, uint64_t* entries) {
    const uint32_t t = (uint32_t)(factorial(n) / factorial(9));
    const uint32_t sl[4] = {0u, 9u, 81u, 585u};
    *tiles = t;
    *slots = sl[n - 9];
    *entries = (uint64_t)t * (uint64_t)(n - 9) * factorial(9);
}

uint64_t footprint_bytes(const ps_puzzle* p, const ps_bfs_options* o, const Geometry& g, int max_layers) {
    uint64_t arrays = (o->bits == 2 ? 1 : o->fused_layers ? 4 : 3) * g.nwords * 8;
    if (o->bits == 3 && o->cross_marks == PS_CROSS_BYTES) arrays += g.nwords * 64;
    if (o->bits == 3 && o->cross_marks == PS_CROSS_LISTS && p->kind == PS_PANCAKE && p->size >= 10 && p->size <= 12) {
        uint32_t tiles, slots;
        uint64_t entries;
        list_geometry(p->size, &tiles, &slots, &entries);
        arrays += entries * 4 + (uint64_t)tiles * slots * 4 + tiles;
        // batched mode: a second parity of lists, counts and flags, tile flags,
        // counts and the list of tiles a launch visits
        if (o->mode == PS_BFS_BATCHED) {
            arrays += entries * 4 + (uint64_t)tiles * slots * 4 + tiles + 2ull * tiles + 12ull * tiles + 4ull * tiles;
            if (o->direction == PS_DIR_AUTO) arrays += g.nwords * 8;  // the visited states
        }
    } else if (o->bits == 3 && o->mode == PS_BFS_BATCHED) {
        // the list of tiles a top-down expand visits: 4 bytes per tile, tiles
        // of at least 64 words unless tile_bits_log2 asks for smaller ones
        const uint64_t tw = o->tile_bits_log2 > 0 && o->tile_bits_log2 < 12 ? 1ull << std::max(o->tile_bits_log2 - 6, 0) : 64;
        arrays += (g.nwords + tw - 1) / tw * 4;
    }
    const uint64_t meta = (o->fused_layers ? 4ull : 3ull) * g.nseg * 8 + ((g.nseg + 31) / 32) * 4 * (o->fused_layers ? 2 : 1) +
                          2ull * max_layers * 8 + sizeof(BfsDevState);
    return arrays + meta;
}

}  // namespace

int enqueue_layers(ps_bfs* b, int n) {
    const bool use_graph = b->opt.mode == PS_BFS_BATCHED && b->opt.graph;
    b->collect_kernel_times = false;
    if (b->fused && !b->primed)
        if (int r = prime_fused(b)) return r;
    if (b->tstep && !b->primed)
        if (int r = prime_tstep(b)) return r;
    if (!use_graph) {
        for (int i = 0; i < n; ++i)
            if (int r = launch_batched_layer(b, false)) return r;
        return PS_OK;
    }
    if (int r = build_graph(b)) return r;
    for (int i = 0; i < n; i += b->graph_layers) {
        PS_HIP_CHECK(hipGraphLaunch(b->graph_exec, b->stream));
        b->stats.expand_launches += b->graph_layers;
        b->stats.finalize_launches += b->bits == 3 && !b->fused && !b->tstep ? b->graph_layers : 0;
        b->stats.apply_launches += b->lists && !b->fused ? b->graph_layers : 0;
    }
    return PS_OK;
}

}  // namespace ps

using namespace ps;

// Reads the device state into the stats (call after the stream is idle).
int ps::collect_stats(ps_bfs* b) {
    PS_HIP_CHECK(hipMemcpyAsync(b->h_st, b->st, sizeof(BfsDevState), hipMemcpyDeviceToHost, b->stream));
    PS_HIP_CHECK(hipStreamSynchronize(b->stream));
    const BfsDevState& st = *b->h_st;
    if (st.error & 1u) return fail(PS_ERR_INDEX_RANGE, "a child index exceeded the index space (%llu states)",
                                   (unsigned long long)b->nbits);
    if (st.error) return fail(PS_ERR_INTERNAL, "cross-tile list out of range (device error bits %u)", st.error);
    b->stats.layers = st.layers_recorded;
    b->stats.expanded = st.expanded;
    b->stats.children = st.children;
    if (b->bits == 3) {
        const uint64_t fin_words_bytes = b->info.has_odd_cycles ? 40 : 32;
        b->stats.algorithmic_bytes =
            8ull * st.scan_words + 8ull * st.children + fin_words_bytes * st.fin_full_words + 8ull * st.fin_clear_words;
        // bottom-up layers: old and cur read per word, one 8-byte read per probe outside the tile
        b->stats.algorithmic_bytes += 16ull * st.pull_words + 8ull * st.pull_probes;
        b->stats.expand_bytes = 8ull * st.scan_words + 8ull * st.children + 16ull * st.pull_words + 8ull * st.pull_probes;
        b->stats.scan_words = st.scan_words;
        b->stats.pull_words = st.pull_words;
        b->stats.pull_probes = st.pull_probes;
    } else {
        // two-bit: every layer reads and writes the whole 2-bit array, 8 bytes per child
        b->stats.algorithmic_bytes = 16ull * b->nwords * (st.layers_recorded) + 8ull * st.children;
        b->stats.expand_bytes = b->stats.algorithmic_bytes;
        b->stats.scan_words = b->nwords * st.layers_recorded;
    }
    double ems = 0, fms = 0;
    for (auto& sp : b->expand_spans) {
        float ms = 0;
        PS_HIP_CHECK(hipEventElapsedTime(&ms, b->ev_pool[sp.first], b->ev_pool[sp.second]));
        ems += ms;
    }
    for (auto& sp : b->fin_spans) {
        float ms = 0;
        PS_HIP_CHECK(hipEventElapsedTime(&ms, b->ev_pool[sp.first], b->ev_pool[sp.second]));
        fms += ms;
    }
    double ams = 0;
    for (auto& sp : b->apply_spans) {
        float ms = 0;
        PS_HIP_CHECK(hipEventElapsedTime(&ms, b->ev_pool[sp.first], b->ev_pool[sp.second]));
        ams += ms;
    }
    double pms = 0;
    for (auto& sp : b->pull_spans) {
        float ms = 0;
        PS_HIP_CHECK(hipEventElapsedTime(&ms, b->ev_pool[sp.first], b->ev_pool[sp.second]));
        pms += ms;
    }
    b->stats.expand_ms = ems;
    b->stats.finalize_ms = fms;
    b->stats.apply_ms = ams;
    b->stats.pull_ms = pms;
    b->stats.list_entries = b->lists ? st.expanded * (uint64_t)(b->puzzle.size - 9) : 0;
    // the fused kernel also finalizes: its bytes are the whole search's but
    // for the bottom-up layers of k_pkl_pull
    if (b->fused) {
        b->stats.pull_bytes = 16ull * st.pull_words + 8ull * st.pull_probes;
        b->stats.expand_bytes = b->stats.algorithmic_bytes - b->stats.pull_bytes;
    }
    return PS_OK;
}

extern "C" {

void ps_bfs_options_default(ps_bfs_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->bits = 3;
    o->mode = PS_BFS_BATCHED;
    o->hash = 1;
    o->graph = 1;
    o->graph_layers = 4;
}

int ps_bfs_footprint(const ps_puzzle* p, const ps_bfs_options* opt, uint64_t* bytes) {
    PS_API_BEGIN
    if (!p || !opt || !bytes) return fail(PS_ERR_INVALID_ARGUMENT, "NULL argument");
    ps_puzzle_info info;
    if (int r = puzzle_query(p, &info)) return r;
    Geometry g;
    if (int r = geometry(p, opt, info, &g)) return r;
    const int ml = opt->max_layers > 0 ? std::min(opt->max_layers, MAX_LAYERS_CAP) : MAX_LAYERS_CAP;
    *bytes = footprint_bytes(p, opt, g, ml);
    return PS_OK;
    PS_API_END
}

int ps_bfs_create(const ps_puzzle* p, const ps_bfs_options* opt_in, ps_stream s, ps_bfs** out) {
    PS_API_BEGIN
    if (int r = require_runtime()) return r;
    if (!p || !out) return fail(PS_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    ps_bfs_options opt;
    if (opt_in) opt = *opt_in;
    else ps_bfs_options_default(&opt);
    if (opt.bits != 2 && opt.bits != 3) return fail(PS_ERR_INVALID_ARGUMENT, "bits must be 2 or 3");
    if (opt.mode != PS_BFS_BATCHED && opt.mode != PS_BFS_PLAIN) return fail(PS_ERR_INVALID_ARGUMENT, "bad mode");
    if (opt.direction < PS_DIR_TOP_DOWN || opt.direction > PS_DIR_BOTTOM_UP)
        return fail(PS_ERR_INVALID_ARGUMENT, "direction must be PS_DIR_TOP_DOWN, PS_DIR_AUTO or PS_DIR_BOTTOM_UP");
    if (opt.cross_marks != PS_CROSS_BYTES && opt.cross_marks != PS_CROSS_ATOMIC && opt.cross_marks != PS_CROSS_LISTS)
        return fail(PS_ERR_INVALID_ARGUMENT, "cross_marks must be PS_CROSS_BYTES, PS_CROSS_ATOMIC or PS_CROSS_LISTS");
    if (opt.bits == 2 && opt.mode == PS_BFS_PLAIN) opt.mode = PS_BFS_BATCHED;
    ps_puzzle_info info;
    if (int r = puzzle_query(p, &info)) return r;
    Geometry g;
    if (int r = geometry(p, &opt, info, &g)) return r;
    if (opt.direction == PS_DIR_BOTTOM_UP && !every_index_a_state(*p, g.nbits))
        return fail(PS_ERR_UNSUPPORTED,
                    "bottom-up layers need a puzzle whose every index is a state (not symmetric Hanoi, "
                    "packed sliding, more than 12 pancakes, or an index_count beyond the last state)");
    if (opt.cross_marks == PS_CROSS_LISTS) {
        if (opt.bits != 3 || p->kind != PS_PANCAKE || p->size < 10 || p->size > 12)
            return fail(PS_ERR_UNSUPPORTED, "PS_CROSS_LISTS covers the three-bit search of 10..12 pancakes");
        if (g.nbits < factorial(p->size))
            return fail(PS_ERR_UNSUPPORTED, "PS_CROSS_LISTS needs index_count >= %d!", p->size);
        if (opt.direction == PS_DIR_BOTTOM_UP || (opt.direction == PS_DIR_AUTO && opt.mode != PS_BFS_BATCHED))
            return fail(PS_ERR_UNSUPPORTED,
                        "PS_CROSS_LISTS runs top-down layers, or with PS_DIR_AUTO in batched mode bottom-up ones "
                        "where the frontier outnumbers the unvisited states");
        if (opt.tile_bits_log2 != 0)
            return fail(PS_ERR_UNSUPPORTED, "PS_CROSS_LISTS uses its own tiles of 9! states (tile_bits_log2 = 0)");
    }
    if (opt.fused_layers != 0 && opt.fused_layers != 1) return fail(PS_ERR_INVALID_ARGUMENT, "fused_layers must be 0 or 1");
    if (opt.fused_layers &&
        (opt.bits != 3 || opt.mode != PS_BFS_BATCHED || opt.cross_marks != PS_CROSS_ATOMIC ||
         opt.direction != PS_DIR_TOP_DOWN || opt.skip_visited || opt.tile_bits_log2 < 0))
        // a tile's expansion reads only its own frontier words and ORs every
        // child outside its window straight into the next array
        return fail(PS_ERR_UNSUPPORTED,
                    "fused_layers covers batched three-bit PS_CROSS_ATOMIC top-down searches with an LDS window "
                    "and without skip_visited");
    // the puzzle must have a device expander
    if (int r = dispatch<FAM_ALL>(*p, PS_EXPAND_FULL, [](auto) { return PS_OK; })) return r;

    ps_bfs* b = new ps_bfs();
    b->puzzle = *p;
    b->opt = opt;
    b->info = info;
    b->bits = opt.bits;
    b->nbits = g.nbits;
    b->nwords = g.nwords;
    b->nseg = g.nseg;
    b->seg_shift = g.seg_shift;
    b->seg_words = g.seg_words;
    b->max_layers = opt.max_layers > 0 ? std::min(opt.max_layers, MAX_LAYERS_CAP) : MAX_LAYERS_CAP;
    b->graph_layers = opt.graph_layers > 0 ? opt.graph_layers : 4;
    auto cleanup = [&](int rc) {
        destroy_bfs(b);
        return rc;
    };
    if (s) {
        b->stream = stream_of(s);
    } else {
        if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
            return cleanup(fail(PS_ERR_HIP, "stream creation failed"));
        b->own_stream = true;
    }
    auto dalloc = [&](void** ptr, uint64_t bytes) -> int {
        hipError_t e = hipMalloc(ptr, bytes ? bytes : 16);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(PS_ERR_OUT_OF_MEMORY, "hipMalloc(%llu) failed: %s", (unsigned long long)bytes,
                        hipGetErrorString(e));
        }
        return PS_OK;
    };
    const int narr = b->bits == 2 ? 1 : 3;
    for (int i = 0; i < narr; ++i)
        if (int r = dalloc((void**)&b->arr[i], b->nwords * 8)) return cleanup(r);
    if (b->bits == 2) b->arr[1] = b->arr[2] = b->arr[0];
    if (b->bits == 3 && opt.cross_marks == PS_CROSS_BYTES)
        if (int r = dalloc((void**)&b->xmark, b->nwords * 64)) return cleanup(r);
    if (b->bits == 3 && opt.cross_marks == PS_CROSS_LISTS) {
        uint64_t entries;
        list_geometry(p->size, &b->xtiles, &b->xslots, &entries);
        b->lists = 1;
        if (int r = dalloc((void**)&b->xlists, entries * 4)) return cleanup(r);
        if (int r = dalloc((void**)&b->xcounts, (uint64_t)b->xtiles * b->xslots * 4)) return cleanup(r);
        if (int r = dalloc((void**)&b->xflags, b->xtiles)) return cleanup(r);
        if (opt.mode == PS_BFS_BATCHED) {
            b->fused = 1;
            if (int r = dalloc((void**)&b->xlists1, entries * 4)) return cleanup(r);
            if (int r = dalloc((void**)&b->xcounts1, (uint64_t)b->xtiles * b->xslots * 4)) return cleanup(r);
            if (int r = dalloc((void**)&b->xflags1, b->xtiles)) return cleanup(r);
            if (int r = dalloc((void**)&b->wflags, 2ull * b->xtiles)) return cleanup(r);
            if (int r = dalloc((void**)&b->tcount, 12ull * b->xtiles)) return cleanup(r);
            if (int r = dalloc((void**)&b->active, 4ull * b->xtiles)) return cleanup(r);
            if (opt.direction == PS_DIR_AUTO)
                if (int r = dalloc((void**)&b->visited, b->nwords * 8)) return cleanup(r);
        }
    }
    if (int r = dalloc((void**)&b->seg_count, (uint64_t)(opt.fused_layers ? 4 : 3) * b->nseg * 8)) return cleanup(r);
    if (int r = dalloc((void**)&b->touched, (uint64_t)((b->nseg + 31) / 32) * 4)) return cleanup(r);
    if (int r = dalloc((void**)&b->layer_counts, (uint64_t)b->max_layers * 8)) return cleanup(r);
    if (int r = dalloc((void**)&b->layer_hashes, (uint64_t)b->max_layers * 8)) return cleanup(r);
    if (int r = dalloc((void**)&b->st, sizeof(BfsDevState))) return cleanup(r);
    if (hipHostMalloc((void**)&b->h_st, sizeof(BfsDevState), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&b->h_seg, (size_t)3 * b->nseg * 8, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&b->h_touched, (size_t)((b->nseg + 31) / 32) * 4 + 4, hipHostMallocDefault) != hipSuccess)
        return cleanup(fail(PS_ERR_OUT_OF_MEMORY, "pinned host allocation failed"));
    std::memset(b->h_st, 0, sizeof(BfsDevState));
    if (hipEventCreate(&b->ev_run0) != hipSuccess || hipEventCreate(&b->ev_run1) != hipSuccess)
        return cleanup(fail(PS_ERR_HIP, "event creation failed"));

    const int cus = hip_cu_count();
    if (b->bits == 2) {
        b->use_window = 0;
        b->tile_words = std::min<uint64_t>(b->nwords, 2048);
        if (b->tile_words == 0) b->tile_words = 1;
        b->ntiles = (b->nwords + b->tile_words - 1) / b->tile_words;
        int occ = 0;
        psk::Bfs2Args a2{};
        LaunchCfg c{256, 0, (size_t)4 * psk::QCAP * 4, b->stream};
        if (int r = bfs2_layer(*p, a2, c, &occ)) return cleanup(r);
        b->grid_bfs2 = (uint32_t)std::max<int64_t>(1, std::min<int64_t>((int64_t)occ * cus, (int64_t)b->ntiles));
    } else {
        // tile (LDS window) geometry
        uint64_t block_bits;
        if (opt.tile_bits_log2 > 0) block_bits = 1ull << std::min(opt.tile_bits_log2, 40);
        else block_bits = puzzle_locality_block(p);
        b->use_window = opt.tile_bits_log2 >= 0 ? 1 : 0;
        b->pull_capable = (opt.direction != PS_DIR_TOP_DOWN && b->use_window && opt.mode == PS_BFS_BATCHED &&
                           every_index_a_state(*p, g.nbits) && !b->lists) ? 1 : 0;
        if (!b->use_window) block_bits = 1ull << 18;
        if (b->lists) block_bits = factorial(9);
        b->tile_words = std::max<uint64_t>(1, block_bits / 64);
        // the fused list kernel fills a CU with one 1024-thread block: half as many
        // tiles in flight as two 512-thread blocks, so fewer partly written list
        // lines are evicted from L2
        b->tpb_expand = opt.block_threads ? opt.block_threads : (b->lists && opt.mode == PS_BFS_BATCHED ? 1024 : 512);
        if (b->tpb_expand != 256 && b->tpb_expand != 512 && !(b->lists && b->tpb_expand == 1024))
            return cleanup(fail(PS_ERR_INVALID_ARGUMENT,
                                "block_threads must be 256 or 512 (or 1024 with PS_CROSS_LISTS)"));
        const size_t lds_limit = hip_lds_limit();
        while (b->use_window && b->tile_words > 1 &&
               psk::bfs3_lds_bytes(b->tpb_expand, b->tile_words, 1, b->nseg, b->pull_capable) + 2048 > lds_limit)
            b->tile_words /= 2;
        if (b->tile_words > b->nwords) b->tile_words = b->nwords;
        b->ntiles = (b->nwords + b->tile_words - 1) / b->tile_words;
        b->lds_expand = psk::bfs3_lds_bytes(b->tpb_expand, b->tile_words, b->use_window, b->nseg, b->pull_capable);
        int occ = 0;
        if (b->lists) {
            b->tile_words = factorial(9) / 64;
            b->ntiles = (b->nwords + b->tile_words - 1) / b->tile_words;
            b->lds_expand = pkl_expand_lds_bytes(*p, b->tpb_expand, b->nseg);
            LaunchCfg c{b->tpb_expand, 0, b->lds_expand, b->stream};
            if (int r = pkl_expand(*p, psk::ExpandArgs{}, psk::ListArgs{}, c, &occ)) return cleanup(r);
            b->lds_apply = pkl_apply_lds_bytes(*p, b->tpb_apply);
            int occ_a = 0;
            LaunchCfg ca{b->tpb_apply, 0, b->lds_apply, b->stream};
            if (int r = pkl_apply(*p, psk::ApplyArgs{}, ca, &occ_a)) return cleanup(r);
            if (occ_a < 1) return cleanup(fail(PS_ERR_UNSUPPORTED, "list apply kernel does not fit (LDS %zu bytes)", b->lds_apply));
            b->grid_apply = (uint32_t)std::max<int64_t>(1, std::min<int64_t>((int64_t)occ_a * cus, (int64_t)b->xtiles));
            if (b->fused) {
                b->lds_step = pkl_step_lds_bytes(*p, b->tpb_expand);
                int occ_s = 0;
                LaunchCfg cs{b->tpb_expand, 0, b->lds_step, b->stream};
                if (int r = pkl_step(*p, psk::StepArgs{}, cs, &occ_s)) return cleanup(r);
                if (occ_s < 1) return cleanup(fail(PS_ERR_UNSUPPORTED, "list step kernel does not fit (LDS %zu bytes)", b->lds_step));
                const int64_t want_s = opt.grid_blocks > 0 ? (int64_t)opt.grid_blocks : (int64_t)occ_s * cus;
                b->grid_step = (uint32_t)std::max<int64_t>(1, std::min<int64_t>(want_s, (int64_t)b->xtiles));
                if (b->visited) {
                    if (b->tpb_expand != 512 && b->tpb_expand != 1024)
                        return cleanup(fail(PS_ERR_UNSUPPORTED, "bottom-up list layers run 512 or 1024 threads"));
                    b->lds_pull = pkl_pull_lds_bytes(*p, b->tpb_expand);
                    int occ_p = 0;
                    LaunchCfg cp{b->tpb_expand, 0, b->lds_pull, b->stream};
                    if (int r = pkl_pull(*p, psk::PullArgs{}, cp, &occ_p)) return cleanup(r);
                    if (occ_p < 1)
                        return cleanup(fail(PS_ERR_UNSUPPORTED, "bottom-up list kernel does not fit (LDS %zu bytes)", b->lds_pull));
                    b->grid_pull = (uint32_t)std::max<int64_t>(1, std::min<int64_t>((int64_t)occ_p * cus, (int64_t)b->xtiles));
                }
            }
        } else {
            LaunchCfg c{b->tpb_expand, 0, b->lds_expand, b->stream};
            if (int r = bfs3_expand(*p, psk::ExpandArgs{}, c, &occ)) return cleanup(r);
        }
        if (occ < 1) return cleanup(fail(PS_ERR_UNSUPPORTED, "expand kernel does not fit (LDS %zu bytes)", b->lds_expand));
        const int64_t want = opt.grid_blocks > 0 ? (int64_t)opt.grid_blocks : (int64_t)occ * cus;
        const int64_t tiles_live = b->lists ? (int64_t)b->xtiles : (int64_t)b->ntiles;
        b->grid_expand = (uint32_t)std::max<int64_t>(1, std::min<int64_t>(want, tiles_live));
        b->chunks_per_seg = (b->seg_words + psk::FIN_CHUNK_WORDS - 1) / psk::FIN_CHUNK_WORDS;
        b->nchunks = (uint64_t)b->nseg * b->chunks_per_seg;
        b->grid_fin = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)cus * 8, b->nchunks));
        if (opt.mode == PS_BFS_BATCHED && !b->lists)
            if (int r = dalloc((void**)&b->tile_list, b->ntiles * 4)) return cleanup(r);
        if (opt.fused_layers) {
            if (b->nseg > 1 && b->seg_words % b->tile_words != 0)
                return cleanup(fail(PS_ERR_INVALID_ARGUMENT,
                                    "fused_layers needs tiles inside one segment (%llu-word tiles, %llu-word segments)",
                                    (unsigned long long)b->tile_words, (unsigned long long)b->seg_words));
            b->tstep = 1;
            if (int r = dalloc((void**)&b->arr[3], b->nwords * 8)) return cleanup(r);
            if (int r = dalloc((void**)&b->touched1, (uint64_t)((b->nseg + 31) / 32) * 4)) return cleanup(r);
            int occ_t = 0;
            LaunchCfg ct{b->tpb_expand, 0, b->lds_expand, b->stream};
            if (int r = bfs3_tile_step(*p, psk::ExpandArgs{}, psk::TileStepArgs{}, ct, &occ_t)) return cleanup(r);
            if (occ_t < 1) return cleanup(fail(PS_ERR_UNSUPPORTED, "tile step kernel does not fit (LDS %zu bytes)", b->lds_expand));
            const int64_t want_t = opt.grid_blocks > 0 ? (int64_t)opt.grid_blocks : (int64_t)occ_t * cus;
            b->grid_tstep = (uint32_t)std::max<int64_t>(1, std::min<int64_t>(want_t, (int64_t)b->ntiles));
        }
    }
    *out = b;
    return PS_OK;
    PS_API_END
}

int ps_bfs_destroy(ps_bfs* b) {
    PS_API_BEGIN
    if (!b) return PS_OK;
    if (runtime_ready()) (void)hipStreamSynchronize(b->stream);
    return destroy_bfs(b);
    PS_API_END
}

int ps_bfs_reset(ps_bfs* b, uint64_t start) {
    PS_API_BEGIN
    if (!b) return fail(PS_ERR_INVALID_ARGUMENT, "NULL bfs");
    if (start >= b->nbits)
        return fail(PS_ERR_INDEX_RANGE, "start index %llu outside [0, %llu)", (unsigned long long)start,
                    (unsigned long long)b->nbits);
    const int narr = b->bits == 2 ? 1 : 3;
    for (int i = 0; i < narr; ++i)
        if (int r = launch_fill(b->arr[i], 0, b->nwords * 8, b->stream)) return r;
    if (b->xmark)
        if (int r = launch_fill(b->xmark, 0, b->nwords * 64, b->stream)) return r;
    if (b->xcounts)
        if (int r = launch_fill(b->xcounts, 0, (uint64_t)b->xtiles * b->xslots * 4, b->stream)) return r;
    if (b->xflags)
        if (int r = launch_fill(b->xflags, 0, b->xtiles, b->stream)) return r;
    if (b->fused) {
        if (int r = launch_fill(b->xcounts1, 0, (uint64_t)b->xtiles * b->xslots * 4, b->stream)) return r;
        if (int r = launch_fill(b->xflags1, 0, b->xtiles, b->stream)) return r;
        if (int r = launch_fill(b->wflags, 0, 2ull * b->xtiles, b->stream)) return r;
        if (int r = launch_fill(b->tcount, 0, 12ull * b->xtiles, b->stream)) return r;
        if (b->visited)
            if (int r = launch_fill(b->visited, 0, b->nwords * 8, b->stream)) return r;
    }
    if (b->tstep) {
        if (int r = launch_fill(b->arr[3], 0, b->nwords * 8, b->stream)) return r;
        if (int r = launch_fill(b->touched1, 0, (uint64_t)((b->nseg + 31) / 32) * 4, b->stream)) return r;
    }
    b->primed = false;
    if (int r = launch_fill(b->seg_count, 0, (uint64_t)(b->tstep ? 4 : 3) * b->nseg * 8, b->stream)) return r;
    if (int r = launch_fill(b->touched, 0, (uint64_t)((b->nseg + 31) / 32) * 4, b->stream)) return r;
    // start state, layer 0
    uint64_t word_idx, word;
    if (b->bits == 2) {
        word_idx = start / 32;
        word = 2ull << (2 * (start % 32));  // CUR of layer 0
    } else {
        word_idx = start / 64;
        word = 1ull << (start % 64);
    }
    ResetStart rs{};
    rs.word_ptr = b->arr[0] + word_idx;
    rs.word = word;
    if (b->bits == 3) rs.seg_count = b->seg_count + (b->nseg > 1 ? (start >> b->seg_shift) : 0);
    rs.layer_counts = b->layer_counts;
    if (b->visited) rs.visited_word = b->visited + word_idx;
    rs.layer_hashes = b->layer_hashes;
    rs.h0 = host_word_hash(word_idx, word);
    BfsDevState st{};
    if (b->tile_list) {  // layer 0 visits the start state's tile
        rs.tile_list = b->tile_list;
        rs.tile = (uint32_t)(start / (b->tile_words * 64));
        st.active_tiles = 1;
    }
    st.max_layers = (uint32_t)b->max_layers;
    st.layers_recorded = 1;
    st.visited = 1;
    st.direction = (b->pull_capable || b->visited) ? (uint32_t)b->opt.direction : 0u;
    st.pull = st.direction == PS_DIR_BOTTOM_UP;
    st.done = b->max_layers <= 1 ? 1 : 0;
    rs.st = b->st;
    rs.st_value = st;
    hipLaunchKernelGGL(k_bfs_reset_start, dim3(1), dim3(64), 0, b->stream, rs);
    PS_HIP_CHECK(hipGetLastError());
    PS_HIP_CHECK(hipStreamSynchronize(b->stream));
    *b->h_st = st;
    b->stats = ps_bfs_stats{};
    b->ev_used = 0;
    b->expand_spans.clear();
    b->fin_spans.clear();
    b->apply_spans.clear();
    b->pull_spans.clear();
    b->layers_done = 1;
    b->start_index = start;
    return PS_OK;
    PS_API_END
}


int ps_bfs_step(ps_bfs* b, uint64_t* count, uint64_t* hash) {
    PS_API_BEGIN
    if (!b) return fail(PS_ERR_INVALID_ARGUMENT, "NULL bfs");
    if (count) *count = 0;
    if (hash) *hash = 0;
    bool done = false;
    if (b->opt.mode == PS_BFS_PLAIN && b->bits == 3) {
        if (int r = run_plain_layer(b, &done)) return r;
    } else {
        PS_HIP_CHECK(hipMemcpyAsync(b->h_st, b->st, sizeof(BfsDevState), hipMemcpyDeviceToHost, b->stream));
        PS_HIP_CHECK(hipStreamSynchronize(b->stream));
        done = b->h_st->done != 0;
        if (!done)
            if (int r = launch_batched_layer(b, true)) return r;
    }
    PS_HIP_CHECK(hipStreamSynchronize(b->stream));
    if (int r = collect_stats(b)) return r;
    const uint64_t rec = b->h_st->layers_recorded;
    if (done || (int)rec <= b->layers_done) return PS_OK;
    b->layers_done = (int)rec;
    unsigned long long c = 0, h = 0;
    PS_HIP_CHECK(hipMemcpy(&c, b->layer_counts + rec - 1, 8, hipMemcpyDeviceToHost));
    PS_HIP_CHECK(hipMemcpy(&h, b->layer_hashes + rec - 1, 8, hipMemcpyDeviceToHost));
    if (count) *count = c;
    if (hash) *hash = h;
    return PS_OK;
    PS_API_END
}

int ps_bfs_run(ps_bfs* b, uint64_t start, uint64_t* counts, uint64_t* hashes, int32_t cap, int32_t* n_layers) {
    PS_API_BEGIN
    if (!b) return fail(PS_ERR_INVALID_ARGUMENT, "NULL bfs");
    if (int r = ps_bfs_reset(b, start)) return r;
    const bool use_graph = b->opt.mode == PS_BFS_BATCHED && b->opt.graph;
    b->collect_kernel_times = !use_graph;
    // a fused search expands its start layer before the (captured) layer loop
    if (b->fused && !b->primed)
        if (int r = prime_fused(b)) return r;
    if (b->tstep && !b->primed)
        if (int r = prime_tstep(b)) return r;
    if (use_graph)
        if (int r = build_graph(b)) return r;
    PS_HIP_CHECK(hipEventRecord(b->ev_run0, b->stream));
    bool done = false;
    while (!done) {
        if (b->opt.mode == PS_BFS_PLAIN && b->bits == 3) {
            if (int r = run_plain_layer(b, &done)) return r;
            continue;
        }
        if (use_graph) {
            PS_HIP_CHECK(hipGraphLaunch(b->graph_exec, b->stream));
            b->stats.expand_launches += b->graph_layers;
            b->stats.finalize_launches += b->bits == 3 && !b->fused && !b->tstep ? b->graph_layers : 0;
            b->stats.apply_launches += b->lists && !b->fused ? b->graph_layers : 0;
        } else {
            if (int r = launch_batched_layer(b, true)) return r;
        }
        PS_HIP_CHECK(hipMemcpyAsync(b->h_st, b->st, sizeof(BfsDevState), hipMemcpyDeviceToHost, b->stream));
        PS_HIP_CHECK(hipStreamSynchronize(b->stream));
        done = b->h_st->done != 0;
    }
    PS_HIP_CHECK(hipEventRecord(b->ev_run1, b->stream));
    PS_HIP_CHECK(hipStreamSynchronize(b->stream));
    const uint64_t launches_e = b->stats.expand_launches, launches_f = b->stats.finalize_launches;
    if (int r = collect_stats(b)) return r;
    b->stats.expand_launches = launches_e;
    b->stats.finalize_launches = launches_f;
    float total = 0;
    PS_HIP_CHECK(hipEventElapsedTime(&total, b->ev_run0, b->ev_run1));
    b->stats.total_ms = total;
    const int nl = (int)b->h_st->layers_recorded;
    b->layers_done = nl;
    if (n_layers) *n_layers = nl;
    const int ncopy = std::min(nl, cap);
    if (counts && ncopy > 0)
        PS_HIP_CHECK(hipMemcpy(counts, b->layer_counts, (size_t)ncopy * 8, hipMemcpyDeviceToHost));
    if (hashes && ncopy > 0)
        PS_HIP_CHECK(hipMemcpy(hashes, b->layer_hashes, (size_t)ncopy * 8, hipMemcpyDeviceToHost));
    if (nl > cap) return fail(PS_ERR_CAPACITY, "%d layers do not fit in %d entries", nl, cap);
    return PS_OK;
    PS_API_END
}

int ps_bfs_stats_get(const ps_bfs* b, ps_bfs_stats* out) {
    PS_API_BEGIN
    if (!b || !out) return fail(PS_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = b->stats;
    return PS_OK;
    PS_API_END
}

int ps_bfs_frontier(const ps_bfs* b, const uint64_t** d_words, uint64_t* nwords) {
    PS_API_BEGIN
    if (!b || !d_words || !nwords) return fail(PS_ERR_INVALID_ARGUMENT, "NULL argument");
    if (b->bits == 2) {
        *d_words = b->arr[0];
    } else {
        const uint32_t L = b->h_st->layer;
        *d_words = b->arr[b->tstep ? L % 4 : L % 3];
    }
    *nwords = b->nwords;
    return PS_OK;
    PS_API_END
}

}  // extern "C"
