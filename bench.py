"""Benchmark: complete three-bit BFS of the 12-pancake graph on one MI355X.

A step is one complete search of the workload (12! = 479,001,600 states,
11 prefix-reversal moves, identity start, 15 layers), device-resident from the
start state to the empty layer.  The headline value is expanded states per
second of the batched persistent path, with the options the autotuner picks:
per layer one persistent kernel that finalizes and expands every tile
(k_pkl_step) when the tuner picks cross-tile lists, else one persistent expand
and one persistent finalize kernel, then one bookkeeping kernel; the layer loop
is captured in a hipGraph.  The plain per-segment
path (one launch per live segment per layer, host-driven) is timed beside it on
the same input, and so are the other ways of marking children outside a tile.

The roofline object covers the dominant kernel (k_pkl_step with cross-tile
lists, else the scan-expand-mark kernel k_bfs3_expand): algorithmic bytes per
launch (8 B per frontier word scanned + 8 B per child generated, and for
k_pkl_step also 40 B per finalized word and 8 B per cleared word) over its
average launch duration, timed with HIP events on the stream the kernel runs
on.  cpu_baseline is the CPU oracle (a one-thread
restatement of the reference's InMemoryThreeBitBFS step) timed on a bounded
slice of the densest layer's frontier.

The Poisson serving runs (multi-search engine) report p50/p99 job latency.

    python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--no-cpu] [--no-serving]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_fixture(name):
    path = os.path.join(ROOT, "tests", "golden", "oracle_fixtures.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get(name)


def timed_runs(psa, bfs, start, steps, warmup):
    for _ in range(warmup):
        bfs.run(start)
    psa.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        counts, hashes = bfs.run(start)
    psa.sync()
    dt = (time.perf_counter() - t0) / steps
    return dt, counts, hashes


def cpu_baseline(psa, puzzle, target_seconds=12.0):
    """Oracle expand+mark over a bounded slice of the densest frontier layer."""
    import ctypes

    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    n = puzzle.size
    N = math.factorial(n)
    with psa.Bfs(puzzle, index_count=N, graph=0) as b:
        b.reset(0)
        best, best_layer, layer = 0, 0, 0
        sizes = []
        while True:
            c, _ = b.step()
            if c == 0:
                break
            layer += 1
            sizes.append(c)
            if c > best:
                best, best_layer = c, layer
                frontier = b.frontier()
    op = O.pancake(n)
    nxt = np.zeros_like(frontier)
    L = O.lib()
    children = ctypes.c_uint64(0)
    # consecutive slices of the layer until target_seconds of CPU work are done
    step = 1 << 22
    first, parents, nchildren, dt = 0, 0, 0, 0.0
    while first < N and dt < target_seconds:
        span = min(step, N - first)
        t0 = time.perf_counter()
        parents += L.or_bfs3_expand_range(ctypes.byref(op), frontier, nxt, first, span, ctypes.byref(children))
        dt += time.perf_counter() - t0
        nchildren += children.value
        first += span
    span = first
    children = ctypes.c_uint64(nchildren)
    single = {"value": parents / dt, "unit": "expanded states/s", "cores": 1, "kind": "port",
              "sample": f"first {span} indexes of layer {best_layer} of the {n}-pancake BFS "
                        f"({parents} parents, {children.value} children marked, {dt:.1f} s, 1 thread)"}
    return single, cpu_baseline_parallel(op, frontier, N, best_layer, n)


def cpu_baseline_parallel(op, frontier, N, layer, n, target_seconds=10.0):
    """The same expand+mark on every host core the box grants (at most 16),
    ParallelExec-style (Common/ThreadUtil.h:19-36): threads take 2^22-index
    slices of the layer from a shared counter, each marking into its own next
    array, until target_seconds of wall time."""
    import ctypes
    import threading

    import numpy as np
    import oracle_lib as O
    L = O.lib()
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(16, cores))
    step = 1 << 22
    lock = threading.Lock()
    state = {"next": 0, "parents": 0, "children": 0}
    deadline = time.perf_counter() + target_seconds

    def worker():
        nxt = np.zeros_like(frontier)
        ch = ctypes.c_uint64(0)
        while time.perf_counter() < deadline:
            with lock:
                first = state["next"] % N
                state["next"] += step
            span = min(step, N - first)
            got = L.or_bfs3_expand_range(ctypes.byref(op), frontier, nxt, first, span, ctypes.byref(ch))
            with lock:
                state["parents"] += got
                state["children"] += ch.value

    t0 = time.perf_counter()
    threads = [threading.Thread(target=worker) for _ in range(cores)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    dt = time.perf_counter() - t0
    return {"value": state["parents"] / dt, "unit": "expanded states/s", "cores": cores, "kind": "port",
            "sample": f"2^22-index slices of layer {layer} of the {n}-pancake BFS on {cores} threads "
                      f"({state['parents']} parents, {state['children']} children marked, {dt:.1f} s)"}


def copy_calibration(psa, nbytes, reps=5):
    """A plain copy kernel of known size: its FETCH_SIZE / WRITE_SIZE readings
    per dispatch give the scale between those counters and bytes moved."""
    src = psa.DeviceBuffer(nbytes)
    dst = psa.DeviceBuffer(nbytes)
    src.zero()
    psa.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        psa.check(psa.lib().ps_copy_async(dst.ptr, src.ptr, nbytes, None))
    psa.sync()
    dt = (time.perf_counter() - t0) / reps
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "reps": reps,
                      "ms_per_copy": dt * 1e3, "GBps": 2 * nbytes / dt / 1e9}))


def traffic_from_profile(name):
    """HBM bytes per launch from a committed counter summary (see its header
    for the rocprofv3 commands and the copy-kernel scale factors), or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def segment_step_bench(psa, steps, ppm=200000, seed=1234):
    """One 15-puzzle disk segment (2^32 packed ids, blank fixed) expanded from a
    seeded Bernoulli(p) frontier: in-segment marks, op-flag scan, frontier
    stream, cross-segment buckets and their compression."""
    from puzzlesearch_amd import workloads as W
    from puzzlesearch_amd.segment import FsSegment
    seg, blank = W.fifteen_segment(seed)
    t0 = time.perf_counter()
    idx, ops = W.bernoulli_frontier(seed, ppm, blank)
    log(f"segment step: frontier of {len(idx)} ids generated in {time.perf_counter() - t0:.1f} s")
    d_idx = psa.DeviceBuffer.from_array(idx)
    d_ops = psa.DeviceBuffer.from_array(ops)
    with FsSegment(psa.sliding_packed(4, 4), 32) as w:
        w.step_device(seg, d_idx.ptr, d_ops.ptr, len(idx))  # warm-up (grows the buffers)
        dev, wall, res = [], [], None
        for _ in range(steps):
            res = w.step_device(seg, d_idx.ptr, d_ops.ptr, len(idx))
            dev.append(res.device_ms)
            wall.append(res.wall_ms)
        r = res.as_dict()
        parity = None
        fx_path = os.path.join(ROOT, "tests", "golden", "segment_fixtures.json")
        if os.path.exists(fx_path):
            with open(fx_path) as f:
                fx = json.load(f).get(f"fifteen_{ppm}")
            if fx and fx["seed"] == seed:
                # the timed step's outputs against digests of the oracle step on the same frontier
                import hashlib

                def sha(a):
                    return hashlib.sha256(memoryview(np.ascontiguousarray(a)).cast("B")).hexdigest()
                o = w.outputs(res)
                got = (r["count"], sha(o["fr_idx"]), sha(o["fr_ops"]), sha(o["fr_bytes"]),
                       [[a, s, n, len(b), sha(b)] for a, s, n, b in o["buckets"]])
                want = (fx["count"], fx["fr_idx"], fx["fr_ops"], fx["fr_bytes"], fx["buckets"])
                assert got == want, "segment step outputs differ from the oracle step"
                parity = True
        # the same frontier as the disk loop stores it: StreamVInt chunks of
        # 2^18 ids with the op bits as 4-bit bounds; decoded on the device,
        # then the same step (checked against the first run's outputs)
        lib = psa.lib()
        nf = len(idx)
        per = 1 << 18
        nch = (nf + per - 1) // per
        starts = np.minimum(np.arange(nch + 1, dtype=np.uint64) * np.uint64(per), np.uint64(nf))
        d_st = psa.DeviceBuffer.from_array(starts)
        cap = 4 * nch + 17 * nf // 4 + nf // 2 + 4096
        d_stream = psa.DeviceBuffer(cap)
        d_soff = psa.DeviceBuffer(8 * (nch + 1))
        psa.check(lib.ps_svint_encode(d_idx.ptr, d_ops.ptr, d_st.ptr, nch, d_stream.ptr, cap, d_soff.ptr, None))
        stream_bytes = int(d_soff.download(np.uint64, nch + 1)[-1])
        d_idx2 = psa.DeviceBuffer(4 * nf + 64)
        d_ops2 = psa.DeviceBuffer(nf + 64)
        d_vs = psa.DeviceBuffer(8 * (nch + 1))
        ref_count = r["count"]
        swall = []
        for _ in range(steps):
            psa.sync()
            t0 = time.perf_counter()
            psa.check(lib.ps_svint_decode(d_stream.ptr, d_soff.ptr, nch, 1, d_idx2.ptr, d_ops2.ptr, nf, d_vs.ptr,
                                          None))
            res2 = w.step_device(seg, d_idx2.ptr, d_ops2.ptr, nf)
            swall.append((time.perf_counter() - t0) * 1e3)
        assert res2.as_dict()["count"] == ref_count
        if parity:
            o2 = w.outputs(res2)
            assert (sha(o2["fr_idx"]), sha(o2["fr_ops"]), sha(o2["fr_bytes"])) == (fx["fr_idx"], fx["fr_ops"],
                                                                                   fx["fr_bytes"])
    ms = sum(wall) / len(wall)
    dms = sum(dev) / len(dev)
    sms = sum(swall) / len(swall)
    return {"workload": f"15-puzzle segment {seg}, blank cell {blank}, frontier Bernoulli(p={ppm / 1e6}) "
                        f"over {W.FIFTEEN_M} ids, seed {seed}",
            "parents": len(idx), "value": len(idx) / (ms * 1e-3), "unit": "expanded states/s",
            "ms_per_step": ms, "device_ms_per_step": dms,
            "algorithmic_bytes": r["algorithmic_bytes"],
            "hbm_frac_device": r["algorithmic_bytes"] / (dms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "count": r["count"], "frontier_count": r["frontier_count"], "frontier_bytes": r["frontier_bytes"],
            "children_in": r["children_in"], "children_cross": r["children_cross"], "buckets": r["buckets"],
            "cross_bytes": r["cross_bytes"], "parity_with_oracle_step": parity,
            "from_stream": {"input": f"StreamVInt chunks of 2^18 ids with 4-bit op bounds, {stream_bytes} bytes",
                            "ms_per_step": sms, "value": len(idx) / (sms * 1e-3),
                            "decode_and_step_parity": parity}}


def seg3_step_bench(psa, steps, ppm=200000, seed=1234, old_seed=4321):
    """The three-bit BFS segment step (DB_Opt3BitBFS_Solver::Expand) on the same
    15-puzzle segment: current frontier Bernoulli(p) of `seed`, old frontier
    Bernoulli(p) of `old_seed` without the current ids, no incoming marks.
    Marks into the segment's NextArray, the old frontier cleared, the scan,
    the FrontierCompression stream and the cross-segment buckets."""
    from puzzlesearch_amd import workloads as W
    from puzzlesearch_amd.seg3 import Seg3
    seg, blank = W.fifteen_segment(seed)
    cur, _ = W.bernoulli_frontier(seed, ppm, blank)
    old0, _ = W.bernoulli_frontier(old_seed, ppm, blank)
    old = np.setdiff1d(old0, cur, assume_unique=True).astype(np.uint32)
    d_cur = psa.DeviceBuffer.from_array(cur)
    d_old = psa.DeviceBuffer.from_array(old)
    with Seg3(psa.sliding_packed(4, 4), 32) as w:
        w.step_device(seg, d_cur.ptr, len(cur), d_old.ptr, len(old))  # warm-up (grows the buffers)
        dev, wall, res = [], [], None
        for _ in range(steps):
            res = w.step_device(seg, d_cur.ptr, len(cur), d_old.ptr, len(old))
            dev.append(res.device_ms)
            wall.append(res.wall_ms)
        r = res.as_dict()
        parity = None
        fx_path = os.path.join(ROOT, "tests", "golden", "segment_fixtures.json")
        if os.path.exists(fx_path):
            with open(fx_path) as f:
                fx = json.load(f).get(f"seg3_fifteen_{ppm}")
            if fx and fx["seed"] == seed and fx["old_seed"] == old_seed:
                # the timed step's outputs against digests of the oracle step on the same frontiers
                import hashlib

                def sha(a):
                    return hashlib.sha256(memoryview(np.ascontiguousarray(a)).cast("B")).hexdigest()
                o = w.outputs(res)
                got = (r["count"], sha(o["fr_idx"]), sha(o["fr_bytes"]),
                       [[a, sg, n, len(b), sha(b)] for a, sg, n, b in o["buckets"]])
                want = (fx["count"], fx["fr_idx"], fx["fr_bytes"], fx["buckets"])
                assert got == want, "three-bit segment step outputs differ from the oracle step"
                parity = True
    ms = sum(wall) / len(wall)
    dms = sum(dev) / len(dev)
    return {"workload": f"15-puzzle segment {seg}, blank cell {blank}, current frontier Bernoulli(p={ppm / 1e6}) "
                        f"of seed {seed}, old frontier of seed {old_seed}, over {W.FIFTEEN_M} ids",
            "parents": len(cur), "old": len(old), "value": len(cur) / (ms * 1e-3), "unit": "expanded states/s",
            "ms_per_step": ms, "device_ms_per_step": dms, "algorithmic_bytes": r["algorithmic_bytes"],
            "hbm_frac_device": r["algorithmic_bytes"] / (dms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "count": r["count"], "frontier_bytes": r["frontier_bytes"], "children_in": r["children_in"],
            "children_cross": r["children_cross"], "buckets": r["buckets"], "cross_bytes": r["cross_bytes"],
            "parity_with_oracle_step": parity}


TUNE_CANDIDATES = {"segment_bits": [22, 20, 24, 26], "block_threads": [512, 256], "grid_blocks": [0, 256, 1024],
                   "graph_layers": [4, 8, 16], "cross_marks": [0, 1], "tile_bits_log2": [0, 16],
                   "skip_visited": [0], "direction": [0, 1], "fused_layers": [0, 1]}
# 10..12 pancakes: cross-tile lists first (bottom-up and other windows do not combine with them)
PANCAKE_CANDIDATES = dict(TUNE_CANDIDATES, cross_marks=[2, 0, 1], block_threads=[0, 512, 256])


def kernel_rooflines(psa, opts, st):
    """Per-kernel achieved bandwidth of an eager (event-timed) run: batched
    searches with cross-tile lists run the fused gather/finalize/expand kernel
    k_pkl_step per layer and, in layers the device expands bottom-up, k_pkl_pull
    after a finalize-only k_pkl_step.  Each kernel's algorithmic bytes (the
    per-unit figures of the design notes times the units it processed) over
    its summed HIP-event time.  Returns {kernel: {...}}, dominant first."""
    out = {}
    main = ("k_pkl_step" if opts.get("cross_marks") == psa.CROSS_LISTS else
            "k_bfs3_tile_step" if opts.get("fused_layers") else "k_bfs3_expand")
    if st.expand_launches:
        out[main] = {"launches": st.expand_launches, "ms_total": st.expand_ms, "bytes_total": st.expand_bytes}
    if st.pull_launches:
        out["k_pkl_pull"] = {"launches": st.pull_launches, "ms_total": st.pull_ms, "bytes_total": st.pull_bytes}
    for k, v in out.items():
        v["avg_launch_ms"] = v["ms_total"] / v["launches"]
        v["algorithmic_bytes_per_launch"] = v["bytes_total"] / v["launches"]
        v["achieved_GBps"] = v["bytes_total"] / (v["ms_total"] * 1e-3) / 1e9 if v["ms_total"] > 0 else 0.0
    return dict(sorted(out.items(), key=lambda kv: -kv[1]["ms_total"]))


def tuned(psa, p, start, index_count=0, candidates=None):
    """Autotuned options (coordinate descent) for one search."""
    from puzzlesearch_amd import autotune as AT
    t0 = time.perf_counter()
    res = AT.tune(p, start, index_count=index_count, passes=1, runs=2, candidates=candidates or TUNE_CANDIDATES)
    log(f"autotune {p}: {res['options']} -> {res['ms']:.2f} ms ({res['tried']} configurations, "
        f"{time.perf_counter() - t0:.1f} s)")
    return res


# options the autotuner picks for the config-3 searches on MI355X, fixed for
# rocprofv3 runs (--profile-set) so that a profile does not depend on a tuning pass
PROFILE_SETS = {
    "sliding4x3": (lambda psa: psa.sliding(4, 3),
                   dict(block_threads=512, segment_bits=20, cross_marks=1, direction=1)),
    "hanoi16": (lambda psa: psa.hanoi(16), dict(block_threads=512, segment_bits=20, cross_marks=1, fused_layers=1)),
    # the same search with split expand and finalize launches
    "hanoi16_split": (lambda psa: psa.hanoi(16), dict(block_threads=512, segment_bits=20, cross_marks=1)),
}


def profile_set(psa, name, runs):
    """Eager batched searches of one config-3 workload (every kernel on the
    search's stream, no graph), for kernel traces and counters."""
    make, opts = PROFILE_SETS[name]
    with psa.Bfs(make(psa), graph=0, **opts) as b:
        for _ in range(runs):
            t0 = time.perf_counter()
            counts, _ = b.run(0)
            dt = time.perf_counter() - t0
        st = b.stats()
    return {"profile": name, "options": opts, "layers": len(counts), "states": sum(counts), "ms_last_run": dt * 1e3,
            "expand_launches": st.expand_launches, "expand_ms": st.expand_ms, "finalize_ms": st.finalize_ms,
            "children": st.children, "algorithmic_bytes": st.algorithmic_bytes}


def codec_bench(psa, n_chunks=256, reps=5, seed=7):
    """Device-resident codec throughput on frontier-like chunks of 2^18 sorted
    ids (the reference's StreamVInt chunk size): StreamVInt with 4-bit bounds
    (gaps of mean 16, and of mean 160 for tuples of mixed lengths),
    FrontierCompression (gaps of mean 16: StreamVInt chunks; mean 3: bitmap
    chunks) and OpBits4.  Each figure is bytes in + bytes out over the call's
    wall time (the encoders read chunk sizes back to the host once per call);
    every round trip is checked."""
    lib = psa.lib()
    rng = np.random.default_rng(seed)
    per = 1 << 18
    n = n_chunks * per
    out = {"values": n, "chunks": n_chunks}

    def frontier(mean_gap):
        gaps = rng.geometric(1.0 / mean_gap, size=(n_chunks, per)).astype(np.uint64)
        v = np.cumsum(gaps, axis=1) + (np.arange(n_chunks, dtype=np.uint64) << np.uint64(32 - 8))[:, None]
        return v.astype(np.uint32).reshape(-1)

    def timed(fn):
        fn()
        psa.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        psa.sync()
        return (time.perf_counter() - t0) / reps

    starts = (np.arange(n_chunks + 1, dtype=np.uint64) * np.uint64(per))
    d_st = psa.DeviceBuffer.from_array(starts)
    cap = 4 * n_chunks + 17 * n // 4 + n // 2 + 4096
    d_out = psa.DeviceBuffer(cap)
    d_off = psa.DeviceBuffer(8 * (n_chunks + 1))
    d_dec = psa.DeviceBuffer(4 * n + 64)
    d_decb = psa.DeviceBuffer(n + 64)
    d_vs = psa.DeviceBuffer(8 * (n_chunks + 1))

    # StreamVInt with bounds
    vals = frontier(16)
    bounds = rng.integers(0, 16, size=n, dtype=np.uint8)
    d_v = psa.DeviceBuffer.from_array(vals)
    d_b = psa.DeviceBuffer.from_array(bounds)
    enc = lambda: psa.check(lib.ps_svint_encode(d_v.ptr, d_b.ptr, d_st.ptr, n_chunks, d_out.ptr, cap, d_off.ptr, None))
    t_enc = timed(enc)
    nbytes = int(d_off.download(np.uint64, n_chunks + 1)[-1])
    dec = lambda: psa.check(lib.ps_svint_decode(d_out.ptr, d_off.ptr, n_chunks, 1, d_dec.ptr, d_decb.ptr, n,
                                                d_vs.ptr, None))
    t_dec = timed(dec)
    assert np.array_equal(d_dec.download(np.uint32, n), vals) and np.array_equal(d_decb.download(np.uint8, n), bounds)
    moved = 5 * n + nbytes
    out["svint"] = {"stream_bytes": nbytes, "encode_ms": t_enc * 1e3, "decode_ms": t_dec * 1e3,
                    "encode_GBps": moved / t_enc / 1e9, "decode_GBps": moved / t_dec / 1e9,
                    "encode_hbm_frac": moved / t_enc / 1e9 / HBM_PEAK_GBS,
                    "decode_hbm_frac": moved / t_dec / 1e9 / HBM_PEAK_GBS}

    # StreamVInt with bounds, gaps of mean 160: about a fifth of the deltas take two bytes, so tuples of
    # 5..9 bytes mix and every chunk takes the decoder's descriptor walk (at mean 16 every tuple has
    # five bytes and the decoder indexes them directly)
    vals = frontier(160)
    d_v.upload(vals)
    enc = lambda: psa.check(lib.ps_svint_encode(d_v.ptr, d_b.ptr, d_st.ptr, n_chunks, d_out.ptr, cap, d_off.ptr, None))
    t_enc = timed(enc)
    nbytes = int(d_off.download(np.uint64, n_chunks + 1)[-1])
    t_dec = timed(dec)
    assert np.array_equal(d_dec.download(np.uint32, n), vals) and np.array_equal(d_decb.download(np.uint8, n), bounds)
    moved = 5 * n + nbytes
    out["svint_mixed"] = {"stream_bytes": nbytes, "encode_ms": t_enc * 1e3, "decode_ms": t_dec * 1e3,
                          "encode_GBps": moved / t_enc / 1e9, "decode_GBps": moved / t_dec / 1e9}

    # FrontierCompression, StreamVInt-coded (mean gap 16) and bitmap-coded (mean gap 3) chunks
    for name, gap in (("fc_svint", 16), ("fc_bitmap", 3)):
        vals = frontier(gap)
        d_v.upload(vals)
        enc = lambda: psa.check(lib.ps_fc_encode(d_v.ptr, d_st.ptr, n_chunks, 1, d_out.ptr, cap, d_off.ptr, None))
        t_enc = timed(enc)
        nbytes = int(d_off.download(np.uint64, n_chunks + 1)[-1])
        dec = lambda: psa.check(lib.ps_fc_decode(d_out.ptr, d_off.ptr, n_chunks, d_dec.ptr, n, d_vs.ptr, None))
        t_dec = timed(dec)
        assert np.array_equal(d_dec.download(np.uint32, n), vals)
        moved = 4 * n + nbytes
        out[name] = {"stream_bytes": nbytes, "encode_ms": t_enc * 1e3, "decode_ms": t_dec * 1e3,
                     "encode_GBps": moved / t_enc / 1e9, "decode_GBps": moved / t_dec / 1e9}

    # OpBits4: 8 op-flag bytes <-> 4 bytes
    ops = rng.integers(0, 16, size=n, dtype=np.uint8)
    d_ops = psa.DeviceBuffer.from_array(ops)
    enc = lambda: psa.check(lib.ps_opbits_encode(4, d_ops.ptr, n, d_out.ptr, None))
    dec = lambda: psa.check(lib.ps_opbits_decode(4, d_out.ptr, n, d_decb.ptr, None))
    t_enc = timed(enc)
    t_dec = timed(dec)
    assert np.array_equal(d_decb.download(np.uint8, n), ops)
    moved = n + n // 2
    out["opbits4"] = {"encode_ms": t_enc * 1e3, "decode_ms": t_dec * 1e3, "encode_GBps": moved / t_enc / 1e9,
                      "decode_GBps": moved / t_dec / 1e9}
    for buf in (d_st, d_out, d_off, d_dec, d_decb, d_vs, d_v, d_b, d_ops):
        buf.free()
    return out


def batched_vs_plain(psa, name, p, start, steps, fixture=None):
    """Batched persistent (tuned) vs plain per-segment launches of one search;
    both must equal each other and, where an oracle fixture exists, its layer
    counts and frontier hashes."""
    res = tuned(psa, p, start)
    opts = dict(res["options"])
    with psa.Bfs(p, graph=1, **opts) as b:
        dt_b, counts, hashes = timed_runs(psa, b, start, steps, 1)
    popts = {k: v for k, v in opts.items() if k in ("segment_bits", "block_threads", "tile_bits_log2", "cross_marks")}
    with psa.Bfs(p, mode=psa.BFS_PLAIN, **popts) as b:
        dt_p, counts_p, hashes_p = timed_runs(psa, b, start, 1, 1)
        st_p = b.stats()
    assert (counts_p, hashes_p) == (counts, hashes)
    n_states = sum(counts)
    out = {"workload": name, "states": n_states, "layers": len(counts), "tuned_options": opts,
           "batched_ms": dt_b * 1e3, "plain_ms": dt_p * 1e3, "plain_expand_launches": st_p.expand_launches,
           "batched_states_per_s": n_states / dt_b, "batched_vs_plain": dt_p / dt_b}
    if fixture is not None:
        assert counts == fixture["counts"], f"{name}: layer counts differ from the oracle fixture"
        assert [str(h) for h in hashes] == fixture["hashes"], f"{name}: frontier hashes differ from the oracle fixture"
        out["parity_with_oracle_fixture"] = True
    log(f"{name}: batched {dt_b * 1e3:.2f} ms, plain {dt_p * 1e3:.2f} ms ({st_p.expand_launches} launches)")
    return out


def serving_bench(psa, rates=(5.0, 20.0, 50.0), n_jobs=200, seed=2024):
    """Poisson streams of complete-BFS requests over the 16-kind serving mix;
    every result is checked against the same search run alone."""
    from puzzlesearch_amd import serve as S
    ref = {}
    out = []
    for rate in rates:
        res = S.serve(rate=rate, n_jobs=n_jobs, seed=seed)
        exact = True
        for r in res["results"]:
            key = r["kind"]
            if key not in ref:
                name, args = key.split("(")
                nums = [int(x) for x in args.rstrip(")").split(",")]
                spec = (name, nums[0], nums[1] if len(nums) > 1 else 0)
                p, n = S.make_puzzle(*spec)
                with psa.Bfs(p, index_count=n) as b:
                    ref[key] = b.run(0)
            exact &= (r["counts"], r["hashes"]) == ref[key]
        rec = {k: res[k] for k in ("rate", "jobs", "seed", "p50_ms", "p99_ms", "mean_ms", "makespan_s", "states",
                                   "states_per_s", "max_running", "all_complete_and_summed")}
        rec["exact_vs_alone"] = exact
        rec["engine_steps"] = res["engine"]["steps"]
        out.append(rec)
        log(f"serving rate {rate}/s: p50 {res['p50_ms']:.1f} ms, p99 {res['p99_ms']:.1f} ms, "
            f"{res['states_per_s']:.3g} states/s, max running {res['max_running']}, exact={exact}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=12, help="pancake size")
    ap.add_argument("--segment-bits", type=int, default=22)
    ap.add_argument("--grid-blocks", type=int, default=0, help="persistent expand blocks (0: resident blocks x CUs)")
    ap.add_argument("--block-threads", type=int, default=0, help="expand block size (0: default)")
    ap.add_argument("--cross-marks", choices=["auto", "bytes", "atomic", "lists"], default="auto",
                    help="children outside the tile: auto = the tuner's choice (with --profile or --no-tune: "
                         "lists for 10..12 pancakes, else bytes)")
    ap.add_argument("--direction", type=int, default=-1, choices=[-1, 0, 1],
                    help="layer direction without the tuner: 0 top-down, 1 auto (bottom-up where the frontier "
                         "outnumbers the unvisited states); -1: auto with cross-tile lists, else top-down")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-segment-step", action="store_true")
    ap.add_argument("--no-codecs", action="store_true", help="skip the device codec throughput runs")
    ap.add_argument("--profile-codecs", action="store_true", help="only the codec runs (for rocprofv3 runs)")
    ap.add_argument("--no-batched-set", action="store_true", help="skip the 4x3 sliding and Hanoi-16 comparisons")
    ap.add_argument("--no-tune", action="store_true", help="use the default options instead of the autotuner's")
    ap.add_argument("--no-serving", action="store_true", help="skip the Poisson serving runs")
    ap.add_argument("--profile", action="store_true",
                    help="only the event-timed batched search, --steps times (for rocprofv3 runs)")
    ap.add_argument("--profile-set", choices=sorted(PROFILE_SETS),
                    help="only eager batched searches of one config-3 workload, --steps times (for rocprofv3 runs)")
    ap.add_argument("--profile-segment", action="store_true",
                    help="only the 15-puzzle segment step, --steps times (for rocprofv3 runs)")
    ap.add_argument("--profile-seg3", action="store_true",
                    help="only the 15-puzzle three-bit segment step, --steps times (for rocprofv3 runs)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a kernel copy of MIB MiB five times (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    import puzzlesearch_amd as psa
    psa.init(0)
    if args.copy_calibration:
        copy_calibration(psa, args.copy_calibration << 20)
        return
    if args.profile_set:
        print(json.dumps(profile_set(psa, args.profile_set, args.steps)))
        return
    if args.profile_codecs:
        print(json.dumps({"profile": "codecs", **codec_bench(psa)}))
        return
    if args.profile_segment:
        print(json.dumps({"profile": "segment_step", **segment_step_bench(psa, args.steps)}))
        return
    if args.profile_seg3:
        print(json.dumps({"profile": "seg3_step", **seg3_step_bench(psa, args.steps)}))
        return
    props = psa.device_props()
    log("device:", props)
    n = args.n
    N = math.factorial(n)
    p = psa.pancake(n)
    fixture = load_fixture(f"pancake_{n}")

    cm = {"bytes": psa.CROSS_BYTES, "atomic": psa.CROSS_ATOMIC, "lists": psa.CROSS_LISTS}.get(args.cross_marks)
    if cm is None:
        cm = psa.CROSS_LISTS if 10 <= n <= 12 else psa.CROSS_BYTES
    direction = args.direction
    if direction < 0:
        # the tuner's pick for the 12-pancake with lists: bottom-up where the frontier outnumbers the unvisited
        direction = psa.DIR_AUTO if cm == psa.CROSS_LISTS else psa.DIR_TOP_DOWN
    opts = {"segment_bits": args.segment_bits, "cross_marks": cm, "grid_blocks": args.grid_blocks,
            "block_threads": args.block_threads, "direction": direction}
    tune_res = None
    if not args.no_tune and not args.profile and args.cross_marks == "auto":
        tune_res = tuned(psa, p, 0, index_count=N, candidates=PANCAKE_CANDIDATES if 10 <= n <= 12 else None)
        opts = dict(tune_res["options"])
    eager_opts = {k: v for k, v in opts.items() if k != "graph_layers"}

    def eager_roofline(runs):
        # the batched path without the graph, HIP events around every kernel on its stream
        with psa.Bfs(p, index_count=N, graph=0, **eager_opts) as b:
            for _ in range(runs):
                ce, he = b.run(0)
            st = b.stats()
        return ce, he, st, kernel_rooflines(psa, opts, st)

    if args.profile:
        ce, he, st_e, kr = eager_roofline(args.steps)
        print(json.dumps({"profile": "12-pancake batched search, eager", "options": eager_opts,
                          "kernels_last_run": kr, "counts": ce}))
        return

    # batched persistent path, layer loop in a hipGraph
    with psa.Bfs(p, index_count=N, graph=1, **opts) as b:
        dt_b, counts, hashes = timed_runs(psa, b, 0, args.steps, args.warmup)
        st_b = b.stats()
    parity = None
    if fixture:
        parity = counts == fixture["counts"] and [str(h) for h in hashes] == fixture["hashes"]
    log(f"batched: {dt_b * 1e3:.2f} ms/search, layers={len(counts)}, sum={sum(counts)}, parity={parity}")
    assert sum(counts) == N, "layer counts do not cover the state space"
    assert parity is not False, "layer counts or frontier hashes differ from the oracle fixture"

    counts_e, hashes_e, st_e, kr = eager_roofline(2)
    assert counts_e == counts and hashes_e == hashes
    kernel = next(iter(kr))
    dom = kr[kernel]
    exp_launch_ms, exp_bytes_launch, achieved = dom["avg_launch_ms"], dom["algorithmic_bytes_per_launch"], \
        dom["achieved_GBps"]
    for k, v in kr.items():
        log(f"{k}: {v['launches']} launches, {v['ms_total']:.3f} ms total, avg {v['avg_launch_ms']:.3f} ms, "
            f"{v['algorithmic_bytes_per_launch'] / 1e9:.3f} GB/launch -> {v['achieved_GBps']:.0f} GB/s")

    # the same batched path top-down only (every layer expands its frontier,
    # the reference algorithm's work), and with the other ways of marking
    # children outside the tile (top-down)
    other_marks = {}
    top_down_bytes = None
    variants = [("top_down", opts.get("cross_marks"))] if opts.get("direction", 0) != 0 else []
    variants += [("byte_cross_marks", psa.CROSS_BYTES), ("atomic_cross_marks", psa.CROSS_ATOMIC)]
    for name, mode in variants:
        if name != "top_down" and opts.get("cross_marks") == mode:
            continue
        with psa.Bfs(p, index_count=N, graph=1, **dict(opts, cross_marks=mode, direction=psa.DIR_TOP_DOWN)) as b:
            dt_x, counts_x, hashes_x = timed_runs(psa, b, 0, max(1, args.steps // 2), 1)
            st_x = b.stats()
        assert counts_x == counts and hashes_x == hashes
        other_marks[name] = {"value": N / dt_x, "ms_per_step": dt_x * 1e3,
                             "algorithmic_bytes_per_search": st_x.algorithmic_bytes}
        if name == "top_down" or (top_down_bytes is None and opts.get("direction", 0) == 0):
            top_down_bytes = st_x.algorithmic_bytes
        log(f"batched, {name.replace('_', ' ')}: {dt_x * 1e3:.2f} ms/search")
    if top_down_bytes is None:
        top_down_bytes = st_b.algorithmic_bytes

    plain = None
    if not args.no_plain:
        # the plain per-segment path at --segment-bits (the batched list path
        # does not depend on the segment size), and at 2^26, the largest
        # segment of the 2^20..2^26 range and the plain path's fastest
        plain = {}
        for key, sb in (("", args.segment_bits), ("seg26_", 26)):
            # the reference's per-segment loop expands every layer top-down
            popts = dict(eager_opts, segment_bits=sb, direction=psa.DIR_TOP_DOWN)
            with psa.Bfs(p, index_count=N, mode=psa.BFS_PLAIN, **popts) as b:
                dt_p, counts_p, hashes_p = timed_runs(psa, b, 0, max(1, args.steps // 2), 1)
                st_p = b.stats()
            assert counts_p == counts and hashes_p == hashes
            if not key:
                plain.update({"value": N / dt_p, "ms_per_step": dt_p * 1e3, "segment_bits": sb,
                              "expand_launches_per_search": st_p.expand_launches,
                              "finalize_launches_per_search": st_p.finalize_launches})
            else:
                plain.update({key + "value": N / dt_p, key + "ms_per_step": dt_p * 1e3,
                              key + "expand_launches_per_search": st_p.expand_launches})
            log(f"plain, 2^{sb}-state segments: {dt_p * 1e3:.2f} ms/search ({st_p.expand_launches} expand launches)")

    batched_set = None
    if not args.no_batched_set:
        batched_set = [
            batched_vs_plain(psa, "three-bit BFS of the 4x3 sliding-tile puzzle (12! indexes, 239500800 reachable)",
                             psa.sliding(4, 3), 0, max(2, args.steps // 2), load_fixture("sliding_4x3")),
            batched_vs_plain(psa, "three-bit BFS of 4-peg Hanoi with 16 disks (4^16 states)",
                             psa.hanoi(16), 0, 2, load_fixture("hanoi_16")),
        ]

    serving = None
    if not args.no_serving:
        serving = serving_bench(psa)

    codecs = None
    if not args.no_codecs:
        codecs = codec_bench(psa)
        log("codecs:", codecs)

    seg_step = seg3_step = None
    if not args.no_segment_step:
        seg_step = segment_step_bench(psa, max(2, args.steps))
        log("segment step:", seg_step)
        seg3_step = seg3_step_bench(psa, max(2, args.steps))
        log("three-bit segment step:", seg3_step)

    cpu = cpu_par = None
    if not args.no_cpu:
        cpu, cpu_par = cpu_baseline(psa, p)
        log("cpu baseline:", cpu)
        log("cpu baseline, all cores:", cpu_par)

    value = N / dt_b
    traffic = traffic_from_profile(f"{kernel}_pancake{n}_traffic.txt")
    out = {
        "metric": "expanded states/s (batched persistent three-bit BFS, 12-pancake)",
        "value": value,
        "unit": "states/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt_b * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (the complete 12-pancake state space from the identity)",
        "config": {"workload": f"complete three-bit BFS of the {n}-pancake graph ({N} states, "
                               f"{len(counts)} layers)",
                   "layers": len(counts), "segment_bits": opts.get("segment_bits", args.segment_bits),
                   "options": opts, "autotune_configurations": tune_res["tried"] if tune_res else 0,
                   "algorithmic_bytes_per_search": st_b.algorithmic_bytes,
                   "children_per_search": st_b.children},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": kernel, "avg_launch_ms": exp_launch_ms,
                     "algorithmic_bytes_per_launch": exp_bytes_launch},
        "kernels": kr,
        "whole_search_hbm_frac": st_b.algorithmic_bytes / dt_b / 1e9 / HBM_PEAK_GBS,
        # the bytes the top-down (reference) algorithm moves for this search,
        # over the wall time of the search as run
        "top_down_algorithmic_bytes_per_search": top_down_bytes,
        "top_down_bytes_hbm_frac": top_down_bytes / dt_b / 1e9 / HBM_PEAK_GBS,
        **other_marks,
        "plain": plain,
        "batched_vs_plain": (value / plain["value"]) if plain else None,
        "batched_vs_plain_seg26": (value / plain["seg26_value"]) if plain else None,
        "batched_set": batched_set,
        "segment_step": seg_step,
        "seg3_step": seg3_step,
        "codecs": codecs,
        "serving": serving,
        "parity_with_oracle_fixture": parity,
        "cpu_baseline": cpu,
        "cpu_baseline_parallel": cpu_par,
        "device": props["name"],
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
