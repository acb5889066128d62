// The device-resident BFS object shared by the single-search driver
// (bfs.hip) and the multi-search engine (batch.hip).
#pragma once

#include <utility>
#include <vector>

#include "bfs_kernels.h"
#include "ps_internal.h"

using psk::BfsDevState;

namespace ps {
constexpr int BFS_MAX_LAYERS = 1 << 16;
}  // namespace ps

struct ps_bfs {
    ps_puzzle puzzle{};
    ps_bfs_options opt{};
    ps_puzzle_info info{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int bits = 3;
    uint64_t nbits = 0, nwords = 0;
    uint64_t* arr[4] = {nullptr, nullptr, nullptr, nullptr};  // arr[3]: fused_layers only
    uint8_t* xmark = nullptr;  // byte per state, three-bit search with PS_CROSS_BYTES
    uint32_t nseg = 1;
    int seg_shift = 63;
    uint64_t seg_words = 0;
    unsigned long long* seg_count = nullptr;
    uint32_t* touched = nullptr;
    uint32_t* touched1 = nullptr;  // fused_layers: touched of the odd mark layers
    unsigned long long* layer_counts = nullptr;
    unsigned long long* layer_hashes = nullptr;
    BfsDevState* st = nullptr;
    BfsDevState* h_st = nullptr;  // pinned mirror
    unsigned long long* h_seg = nullptr;  // pinned, 3*nseg
    uint32_t* h_touched = nullptr;        // pinned
    int max_layers = ps::BFS_MAX_LAYERS;
    // expand geometry
    int use_window = 1;
    int pull_capable = 0;  // expand LDS also holds the tile's cur words (bottom-up layers)
    uint64_t tile_words = 0, ntiles = 0;
    uint32_t* tile_list = nullptr;  // batched top-down layers: the tiles the next expand visits
    int tpb_expand = 512;
    uint32_t grid_expand = 0;
    size_t lds_expand = 0;
    // finalize geometry
    int tpb_fin = 256;
    uint32_t grid_fin = 0;
    uint64_t chunks_per_seg = 0, nchunks = 0;
    // two-bit geometry
    uint32_t grid_bfs2 = 0;
    // cross-tile child lists (PS_CROSS_LISTS, 10..12 pancakes)
    int lists = 0;
    uint32_t* xlists = nullptr;   // [tile][flip - 10][9!] child ranks
    uint32_t* xcounts = nullptr;  // [tile][slot] fill counts
    uint8_t* xflags = nullptr;    // [tile] a list reaches the tile this layer
    uint32_t xtiles = 0;          // n!/9!
    uint32_t xslots = 0;          // lists per tile
    uint32_t grid_apply = 0;
    int tpb_apply = 1024;
    // batched mode with lists: one fused kernel per layer (k_pkl_step) over
    // lists of two parities; the start layer is expanded by a first launch
    int fused = 0;
    bool primed = false;
    // batched PS_CROSS_ATOMIC with fused_layers: one k_bfs3_tile_step per layer
    int tstep = 0;
    uint32_t grid_tstep = 0;
    uint32_t* xlists1 = nullptr;  // parity 1 (parity 0: xlists, xcounts, xflags)
    uint64_t* visited = nullptr;  // fused lists with PS_DIR_AUTO: every state finalized so far
    uint32_t grid_pull = 0;
    size_t lds_pull = 0;
    uint32_t* xcounts1 = nullptr;
    uint8_t* xflags1 = nullptr;
    uint8_t* wflags = nullptr;    // [2][tile]
    uint32_t* tcount = nullptr;   // [3][tile]
    uint32_t* active = nullptr;   // [tile] the tiles the next step launch visits
    uint32_t grid_step = 0;
    size_t lds_step = 0;
    size_t lds_apply = 0;
    // graph
    hipGraphExec_t graph_exec = nullptr;
    int graph_layers = 4;
    // events for timing
    hipEvent_t ev_run0 = nullptr, ev_run1 = nullptr;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<std::pair<size_t, size_t>> expand_spans, fin_spans, apply_spans, pull_spans;
    bool collect_kernel_times = true;
    ps_bfs_stats stats{};
    int layers_done = 0;
    uint64_t start_index = 0;
};

namespace ps {

// kernel arguments of one search in its current geometry
psk::ExpandArgs expand_args(ps_bfs* b);
psk::FinalizeArgs finalize_args(ps_bfs* b);
psk::LayerEndArgs layer_end_args(ps_bfs* b);
psk::ListArgs list_args(ps_bfs* b);
psk::ApplyArgs apply_args(ps_bfs* b);
psk::StepArgs step_args(ps_bfs* b, bool first);
psk::PullArgs pull_args(ps_bfs* b);
// expands the start layer of a fused search (once after each reset)
int prime_fused(ps_bfs* b);
// expands the start layer of a fused_layers search and lists the first tiles
int prime_tstep(ps_bfs* b);
// reads the device state into b->stats (stream idle)
int collect_stats(ps_bfs* b);
// enqueues one complete layer (expand, finalize, bookkeeping) on b->stream
int launch_batched_layer(ps_bfs* b, bool timed);
// enqueues at least n more layers of a batched search on b->stream, through
// its captured layer graph when it has one (whole replays of graph_layers
// layers); layers after the last one are no-ops on the device
int enqueue_layers(ps_bfs* b, int n);

}  // namespace ps
