// CPU emulation of the multi-search kernels of the fused batch mode
// (k_multi_expand, k_multi_finalize, k_multi_layer_end in
// puzzlesearch_amd/csrc/bfs_kernels.h, with the puzzle table of
// csrc/multi_kinds.h), compiled unchanged for the host over the SIMT emulator
// in tests/cpp/simt, with AddressSanitizer.  A mix of searches runs step by
// step the way csrc/batch.hip drives them:
//   * each job's arrays are allocated with exactly the sizes ps_bfs_create
//     gives them (restated below for the options bfs_opts_for picks in fused
//     mode: three bits, batched, 512-thread expand, byte cross marks, segments
//     of 2^22 states above 2^22 states), so an access past any of them is
//     reported;
//   * the job table is rebuilt as build_table does whenever a job is admitted
//     or retired, and each step launches the three kernels over it;
//   * every barrier and wave operation is checked by the emulator, and every
//     launch has a time limit;
//   * each job's layer counts and frontier hashes are compared with the
//     oracle's InMemoryThreeBitBFS (Common/InMemoryThreeBitBFS.cpp:4-48).
// Usage: multi_emu [--grid G] [--max-steps S] [--segment-bits B] kind:a:b ...   (kind = pancake,
// sliding, hanoi; b = height for sliding).  Exit status 0 = every job equal
// to the oracle.
#include "simt_emu.h"

#include <algorithm>
#include <cinttypes>
#include <memory>
#include <string>
#include <vector>

namespace psk {
alignas(16) thread_local uint32_t smem[simt::LDS_BYTES / 4];
}

#include "multi_kinds.h"
#include "oracle.h"

namespace {

constexpr int MULTI_TPB = 512, FIN_TPB = 256;
constexpr size_t LDS_LIMIT = 163840;  // gfx950 LDS per workgroup
constexpr int MAX_LAYERS = 1 << 16;

uint64_t fact(int n) {
    uint64_t r = 1;
    for (int i = 2; i <= n; ++i) r *= (uint64_t)i;
    return r;
}

uint64_t splitmix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// puzzle_locality_block (csrc/expand.hip), restated
uint64_t locality_block(const ps_puzzle& p, uint64_t n) {
    switch (p.kind) {
    case PS_PANCAKE:
    case PS_SLIDING: {
        const int cells = p.kind == PS_PANCAKE ? p.size : p.size * p.height;
        if (cells >= 11) return fact(9);
        if (cells == 10) return fact(8);
        return 4096;
    }
    case PS_HANOI4:
    case PS_HANOI4_OPT: {
        uint64_t b = 1ull << 18;
        while (b > 4096 && n / b < 256) b >>= 2;
        return b;
    }
    }
    return 4096;
}

struct Job {
    std::string name;
    ps_puzzle p{};
    or_puzzle op{};
    int kind = -1;
    uint64_t nbits = 0, nwords = 0, seg_words = 0, tile_words = 0, ntiles = 0, chunks_per_seg = 0, nchunks = 0;
    uint32_t nseg = 1;
    int seg_shift = 63;
    int odd = 0;
    std::vector<uint64_t> arr[3];
    std::vector<uint8_t> xmark;
    std::vector<unsigned long long> seg_count, layer_counts, layer_hashes;
    std::vector<uint32_t> touched, tile_list;
    std::unique_ptr<psk::BfsDevState> st;
    bool done = false;
    int steps = 0;
};

int g_segment_bits = 22;  // the batch's segment_bits: segments for searches of more states

bool make_job(const std::string& spec, Job* j) {
    char kind[32] = {};
    int a = 0, b = 0;
    if (std::sscanf(spec.c_str(), "%31[a-z]:%d:%d", kind, &a, &b) < 2) return false;
    const std::string k = kind;
    j->name = spec;
    if (k == "pancake") {
        j->p = ps_puzzle{PS_PANCAKE, a, 0, 0};
        j->nbits = fact(a);
    } else if (k == "sliding") {
        j->p = ps_puzzle{PS_SLIDING, a, b, 0};
        j->nbits = fact(a * b);
    } else if (k == "hanoi") {
        j->p = ps_puzzle{PS_HANOI4, a, 0, 0};
        j->nbits = 1ull << (2 * a);
    } else {
        return false;
    }
    j->op = or_puzzle{j->p.kind, j->p.size, j->p.height, j->p.flags};
    j->kind = ps::multi_kind_of(j->p);
    if (j->kind < 0) return false;
    j->odd = or_has_odd_cycles(&j->op);
    // geometry of ps_bfs_create (csrc/bfs.hip) for the fused batch options
    const int segment_bits = (g_segment_bits > 0 && j->nbits > (1ull << g_segment_bits)) ? g_segment_bits : 0;
    j->nwords = (j->nbits + 63) / 64;
    if (segment_bits == 0) {
        j->nseg = 1;
        j->seg_words = j->nwords;
        j->seg_shift = 63;
    } else {
        j->seg_shift = segment_bits;
        j->seg_words = 1ull << (segment_bits - 6);
        j->nseg = (uint32_t)((j->nwords + j->seg_words - 1) / j->seg_words);
    }
    j->tile_words = std::max<uint64_t>(1, locality_block(j->p, j->nbits) / 64);
    while (j->tile_words > 1 && psk::bfs3_lds_bytes(MULTI_TPB, j->tile_words, 1, j->nseg, 0) + 2048 > LDS_LIMIT)
        j->tile_words /= 2;
    if (j->tile_words > j->nwords) j->tile_words = j->nwords;
    j->ntiles = (j->nwords + j->tile_words - 1) / j->tile_words;
    j->chunks_per_seg = (j->seg_words + psk::FIN_CHUNK_WORDS - 1) / psk::FIN_CHUNK_WORDS;
    j->nchunks = (uint64_t)j->nseg * j->chunks_per_seg;
    // the allocations, exact sizes, zeroed as ps_bfs_reset leaves them
    for (auto& v : j->arr) v.assign(j->nwords, 0);
    j->xmark.assign(j->nwords * 64, 0);
    j->seg_count.assign(3ull * j->nseg, 0);
    j->touched.assign((j->nseg + 31) / 32, 0);
    j->layer_counts.assign(MAX_LAYERS, 0);
    j->layer_hashes.assign(MAX_LAYERS, 0);
    j->tile_list.assign(j->ntiles, 0);
    j->st = std::make_unique<psk::BfsDevState>();
    // start state 0, layer 0 (ps_bfs_reset)
    j->arr[0][0] = 1;
    j->seg_count[0] = 1;
    j->layer_counts[0] = 1;
    j->layer_hashes[0] = splitmix(splitmix(0) ^ 1ull);
    j->tile_list[0] = 0;
    psk::BfsDevState st{};
    st.active_tiles = 1;
    st.max_layers = MAX_LAYERS;
    st.layers_recorded = 1;
    st.visited = 1;
    *j->st = st;
    return true;
}

psk::MultiJob table_entry(Job* j, uint32_t slot, uint64_t tile_base, uint64_t chunk_base) {
    psk::MultiJob J{};
    psk::ExpandArgs& a = J.e;  // expand_args
    a.prm = ps::params_of(j->p);
    for (int i = 0; i < 3; ++i) a.bits[i] = j->arr[i].data();
    a.nbits = j->nbits;
    a.nwords = j->nwords;
    a.tile_words = j->tile_words;
    a.tile_begin = 0;
    a.tile_end = j->ntiles;
    a.word_lo = 0;
    a.word_hi = j->nwords;
    a.seg_shift = j->seg_shift;
    a.nseg = j->nseg;
    a.seg_count = j->seg_count.data();
    a.touched = j->touched.data();
    a.st = j->st.get();
    a.persistent = 1;
    a.use_window = 1;
    a.pull_capable = 0;
    a.xmark = j->xmark.data();
    a.skip_visited = 0;
    psk::FinalizeArgs& f = J.f;  // finalize_args
    for (int i = 0; i < 3; ++i) f.bits[i] = j->arr[i].data();
    f.nwords = j->nwords;
    f.seg_words = j->seg_words;
    f.chunks_per_seg = j->chunks_per_seg;
    f.chunk_begin = 0;
    f.chunk_end = j->nchunks;
    f.nseg = j->nseg;
    f.odd = j->odd;
    f.hash = 1;
    f.persistent = 1;
    f.seg_count = j->seg_count.data();
    f.touched = j->touched.data();
    f.st = j->st.get();
    f.xmark = j->xmark.data();
    psk::LayerEndArgs& e = J.l;  // layer_end_args
    e.nbits = j->nbits;
    e.nseg = j->nseg;
    e.seg_count = j->seg_count.data();
    e.touched = j->touched.data();
    e.layer_counts = j->layer_counts.data();
    e.layer_hashes = j->layer_hashes.data();
    e.st = j->st.get();
    e.bits = 3;
    e.active = j->tile_list.data();
    e.ntiles = j->ntiles;
    e.tile_bits = j->tile_words * 64;
    e.seg_shift = j->seg_shift;
    J.kind = (uint32_t)j->kind;
    J.slot = slot;
    J.tile_base = tile_base;
    J.chunk_base = chunk_base;
    return J;
}

}  // namespace

int main(int argc, char** argv) {
    int grid = 3, max_steps = 100000;
    std::vector<std::string> specs;
    for (int i = 1; i < argc; ++i) {
        const std::string s = argv[i];
        if (s == "--grid" && i + 1 < argc) grid = std::atoi(argv[++i]);
        else if (s == "--max-steps" && i + 1 < argc) max_steps = std::atoi(argv[++i]);
        else if (s == "--segment-bits" && i + 1 < argc) g_segment_bits = std::atoi(argv[++i]);
        else specs.push_back(s);
    }
    simt::lds = []() { return reinterpret_cast<unsigned char*>(psk::smem); };
    std::vector<std::unique_ptr<Job>> jobs;
    for (auto& s : specs) {
        auto j = std::make_unique<Job>();
        if (!make_job(s, j.get())) {
            std::fprintf(stderr, "bad job spec %s\n", s.c_str());
            return 2;
        }
        std::printf("job %s: %" PRIu64 " states, %" PRIu64 " words, tile %" PRIu64 " words, %" PRIu64
                    " tiles, %u segments, %" PRIu64 " chunks\n",
                    s.c_str(), j->nbits, j->nwords, j->tile_words, j->ntiles, j->nseg, j->nchunks);
        jobs.push_back(std::move(j));
    }
    std::vector<Job*> running;
    for (auto& j : jobs) running.push_back(j.get());
    std::vector<psk::MultiJob> table;  // exact size: njobs entries
    std::unique_ptr<unsigned long long[]> counters(new unsigned long long[8]());
    std::vector<uint32_t> done_flags;
    bool dirty = true;
    uint64_t total_tiles = 0, total_chunks = 0;
    uint32_t win = 0;
    size_t lds = 0;
    int step = 0, failures = 0;
    long barriers = 0, waveops = 0;
    while (!running.empty()) {
        if (++step > max_steps) {
            std::printf("FAIL: %zu jobs still running after %d steps\n", running.size(), max_steps);
            return 1;
        }
        if (dirty) {  // build_table (csrc/batch.hip)
            table.assign(running.size(), psk::MultiJob{});
            done_flags.assign(running.size(), 0);
            uint64_t tb = 0, cb = 0;
            win = 0;
            for (size_t i = 0; i < running.size(); ++i) {
                table[i] = table_entry(running[i], (uint32_t)i, tb, cb);
                tb += running[i]->ntiles;
                cb += running[i]->nchunks;
                win = std::max<uint32_t>(win, (uint32_t)(running[i]->tile_words * 2));
            }
            total_tiles = tb;
            total_chunks = cb;
            lds = ((size_t)win + (MULTI_TPB / 64) * psk::QCAP + psk::MULTI_TL_WORDS) * 4;
            dirty = false;
        }
        psk::MultiArgs m{};
        m.jobs = table.data();
        m.njobs = (uint32_t)running.size();
        m.total_tiles = total_tiles;
        m.total_chunks = total_chunks;
        m.counters = counters.get();
        m.done = done_flags.data();
        const uint32_t ge = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)grid, total_tiles));
        const uint32_t gf = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)grid, total_chunks));
        const uint32_t w = win;
        const bool trace = std::getenv("EMU_TRACE") != nullptr;
        if (trace)
            std::fprintf(stderr, "step %d: %u jobs, %" PRIu64 " tiles, %" PRIu64 " chunks, lds %zu\n", step, m.njobs,
                         total_tiles, total_chunks, lds);
        auto s1 = simt::launch("k_multi_expand", ge, MULTI_TPB, lds,
                               [&]() { psk::k_multi_expand<ps::MultiKinds, MULTI_TPB>(m, w); }, grid);
        auto s2 = simt::launch("k_multi_finalize", gf, FIN_TPB, 0, [&]() { psk::k_multi_finalize<FIN_TPB>(m); }, grid);
        auto s3 = simt::launch("k_multi_layer_end", m.njobs, 1024, 0, [&]() { psk::k_multi_layer_end<1024>(m); },
                               grid);
        barriers += s1.barriers + s2.barriers + s3.barriers;
        waveops += s1.waveops + s2.waveops + s3.waveops;
        std::vector<Job*> still;
        for (size_t i = 0; i < running.size(); ++i) {
            Job* j = running[i];
            j->steps++;
            if (j->st->error) {
                std::printf("FAIL: job %s: device error bits %u at step %d\n", j->name.c_str(), j->st->error, step);
                return 1;
            }
            if (!done_flags[i]) {
                still.push_back(j);
                continue;
            }
            // retire: compare with the oracle
            const int nl = (int)j->st->layers_recorded;
            std::vector<uint64_t> oc(MAX_LAYERS), oh(MAX_LAYERS);
            const int onl = or_bfs3(&j->op, 0, j->nbits, MAX_LAYERS, oc.data(), oh.data());
            bool ok = onl == nl;
            for (int l = 0; ok && l < nl; ++l) ok = oc[l] == j->layer_counts[l] && oh[l] == j->layer_hashes[l];
            std::printf("%s: job %s done at step %d, %d layers (oracle %d)\n", ok ? "ok" : "FAIL", j->name.c_str(),
                        step, nl, onl);
            if (!ok) {
                ++failures;
                for (int l = 0; l < std::max(nl, onl) && l < 200; ++l)
                    if (l >= nl || l >= onl || oc[l] != j->layer_counts[l] || oh[l] != j->layer_hashes[l])
                        std::printf("  layer %d: %llu vs oracle %" PRIu64 "\n", l,
                                    l < nl ? (unsigned long long)j->layer_counts[l] : 0ull, l < onl ? oc[l] : 0);
            }
            dirty = true;
        }
        running.swap(still);
        std::fflush(stdout);
    }
    std::printf("%d steps, %ld block barriers, %ld wave operations, %d failures\n", step, barriers, waveops, failures);
    return failures ? 1 : 0;
}
