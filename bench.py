"""Benchmark: complete three-bit BFS of the 12-pancake graph on one MI355X.

A step is one complete search of the workload (12! = 479,001,600 states,
11 prefix-reversal moves, identity start, 15 layers), device-resident from the
start state to the empty layer.  The headline value is expanded states per
second of the batched persistent path (one persistent expand kernel, one
persistent finalize kernel and one bookkeeping kernel per layer, the layer loop
captured in a hipGraph).  The plain per-segment path (one launch per live
segment per layer, host-driven) is timed beside it on the same input.

The roofline object covers the dominant kernel (the fused scan-expand-mark
kernel): algorithmic bytes per launch (8 B per frontier word scanned + 8 B per
child generated) over its average launch duration, timed with HIP events on
the stream the kernel runs on.  cpu_baseline is the CPU oracle (a one-thread
restatement of the reference's InMemoryThreeBitBFS step) timed on a bounded
slice of the densest layer's frontier.

    python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--no-cpu]
"""
import argparse
import json
import math
import os
import sys
import time

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
    return {"value": parents / dt, "unit": "expanded states/s", "cores": 1, "kind": "port",
            "sample": f"first {span} indexes of layer {best_layer} of the {n}-pancake BFS "
                      f"({parents} parents, {children.value} children marked, {dt:.1f} s, 1 thread)"}


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
    ms = sum(wall) / len(wall)
    dms = sum(dev) / len(dev)
    return {"workload": f"15-puzzle segment {seg}, blank cell {blank}, frontier Bernoulli(p={ppm / 1e6}) "
                        f"over {W.FIFTEEN_M} ids, seed {seed}",
            "parents": len(idx), "value": len(idx) / (ms * 1e-3), "unit": "expanded states/s",
            "ms_per_step": ms, "device_ms_per_step": dms,
            "algorithmic_bytes": r["algorithmic_bytes"],
            "hbm_frac_device": r["algorithmic_bytes"] / (dms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "count": r["count"], "frontier_count": r["frontier_count"], "frontier_bytes": r["frontier_bytes"],
            "children_in": r["children_in"], "children_cross": r["children_cross"], "buckets": r["buckets"],
            "cross_bytes": r["cross_bytes"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=12, help="pancake size")
    ap.add_argument("--segment-bits", type=int, default=22)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-segment-step", action="store_true")
    ap.add_argument("--profile", action="store_true",
                    help="only the event-timed batched search, --steps times (for rocprofv3 runs)")
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
    props = psa.device_props()
    log("device:", props)
    n = args.n
    N = math.factorial(n)
    p = psa.pancake(n)
    fixture = load_fixture(f"pancake_{n}")

    def eager_roofline(runs):
        # the batched path without the graph, HIP events around every kernel on its stream
        with psa.Bfs(p, index_count=N, segment_bits=args.segment_bits, graph=0) as b:
            for _ in range(runs):
                ce, he = b.run(0)
            st = b.stats()
        ms = st.expand_ms / max(1, st.expand_launches)
        by = st.expand_bytes / max(1, st.expand_launches)
        return ce, he, st, ms, by

    if args.profile:
        ce, he, st_e, ms, by = eager_roofline(args.steps)
        print(json.dumps({"profile": "k_bfs3_expand", "launches_last_run": st_e.expand_launches,
                          "avg_launch_ms": ms, "algorithmic_bytes_per_launch": by,
                          "achieved_GBps": by / (ms * 1e-3) / 1e9, "counts": ce}))
        return

    # batched persistent path, layer loop in a hipGraph
    with psa.Bfs(p, index_count=N, segment_bits=args.segment_bits, graph=1) as b:
        dt_b, counts, hashes = timed_runs(psa, b, 0, args.steps, args.warmup)
        st_b = b.stats()
    parity = None
    if fixture:
        parity = counts == fixture["counts"] and [str(h) for h in hashes] == fixture["hashes"]
    log(f"batched: {dt_b * 1e3:.2f} ms/search, layers={len(counts)}, sum={sum(counts)}, parity={parity}")
    assert sum(counts) == N, "layer counts do not cover the state space"

    counts_e, hashes_e, st_e, exp_launch_ms, exp_bytes_launch = eager_roofline(2)
    assert counts_e == counts and hashes_e == hashes
    achieved = exp_bytes_launch / (exp_launch_ms * 1e-3) / 1e9
    log(f"expand kernel: {st_e.expand_launches} launches, avg {exp_launch_ms:.3f} ms, "
        f"{exp_bytes_launch / 1e9:.3f} GB/launch -> {achieved:.0f} GB/s; finalize {st_e.finalize_ms:.2f} ms total")

    # the same batched path with device-scope atomics for children outside the LDS window
    with psa.Bfs(p, index_count=N, segment_bits=args.segment_bits, graph=1, cross_marks=psa.CROSS_ATOMIC) as b:
        dt_a, counts_a, hashes_a = timed_runs(psa, b, 0, max(1, args.steps // 2), 1)
    assert counts_a == counts and hashes_a == hashes
    log(f"batched, atomic cross marks: {dt_a * 1e3:.2f} ms/search")

    plain = None
    if not args.no_plain:
        with psa.Bfs(p, index_count=N, segment_bits=args.segment_bits, mode=psa.BFS_PLAIN) as b:
            dt_p, counts_p, hashes_p = timed_runs(psa, b, 0, max(1, args.steps // 2), 1)
            st_p = b.stats()
        assert counts_p == counts and hashes_p == hashes
        plain = {"value": N / dt_p, "ms_per_step": dt_p * 1e3,
                 "expand_launches_per_search": st_p.expand_launches,
                 "finalize_launches_per_search": st_p.finalize_launches}
        log(f"plain: {dt_p * 1e3:.2f} ms/search ({st_p.expand_launches} expand launches)")

    seg_step = None
    if not args.no_segment_step:
        seg_step = segment_step_bench(psa, max(2, args.steps))
        log("segment step:", seg_step)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(psa, p)
        log("cpu baseline:", cpu)

    value = N / dt_b
    traffic = traffic_from_profile(f"k_bfs3_expand_pancake{n}_traffic.txt")
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
                               f"{len(counts)} layers), segments of 2^{args.segment_bits} states",
                   "layers": len(counts), "segment_bits": args.segment_bits,
                   "algorithmic_bytes_per_search": st_b.algorithmic_bytes,
                   "children_per_search": st_b.children},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_bfs3_expand", "avg_launch_ms": exp_launch_ms,
                     "algorithmic_bytes_per_launch": exp_bytes_launch},
        "whole_search_hbm_frac": st_b.algorithmic_bytes / dt_b / 1e9 / HBM_PEAK_GBS,
        "atomic_cross_marks": {"value": N / dt_a, "ms_per_step": dt_a * 1e3},
        "plain": plain,
        "batched_vs_plain": (value / plain["value"]) if plain else None,
        "segment_step": seg_step,
        "parity_with_oracle_fixture": parity,
        "cpu_baseline": cpu,
        "device": props["name"],
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
