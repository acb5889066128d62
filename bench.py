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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=12, help="pancake size")
    ap.add_argument("--segment-bits", type=int, default=22)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a kernel copy of MIB MiB five times (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is design-only, see DESIGN.md)"}))
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

    # batched persistent path, layer loop in a hipGraph
    with psa.Bfs(p, index_count=N, segment_bits=args.segment_bits, graph=1) as b:
        dt_b, counts, hashes = timed_runs(psa, b, 0, args.steps, args.warmup)
        st_b = b.stats()
    parity = None
    if fixture:
        parity = counts == fixture["counts"] and [str(h) for h in hashes] == fixture["hashes"]
    log(f"batched: {dt_b * 1e3:.2f} ms/search, layers={len(counts)}, sum={sum(counts)}, parity={parity}")
    assert sum(counts) == N, "layer counts do not cover the state space"

    # the same path without the graph, with events around every kernel (roofline)
    with psa.Bfs(p, index_count=N, segment_bits=args.segment_bits, graph=0) as b:
        b.run(0)
        counts_e, hashes_e = b.run(0)
        st_e = b.stats()
    assert counts_e == counts and hashes_e == hashes
    exp_launch_ms = st_e.expand_ms / max(1, st_e.expand_launches)
    exp_bytes_launch = st_e.expand_bytes / max(1, st_e.expand_launches)
    achieved = exp_bytes_launch / (exp_launch_ms * 1e-3) / 1e9
    log(f"expand kernel: {st_e.expand_launches} launches, avg {exp_launch_ms:.3f} ms, "
        f"{exp_bytes_launch / 1e9:.3f} GB/launch -> {achieved:.0f} GB/s; finalize {st_e.finalize_ms:.2f} ms total")

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

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(psa, p)
        log("cpu baseline:", cpu)

    value = N / dt_b
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
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_bfs3_expand", "avg_launch_ms": exp_launch_ms,
                     "algorithmic_bytes_per_launch": exp_bytes_launch},
        "whole_search_hbm_frac": st_b.algorithmic_bytes / dt_b / 1e9 / HBM_PEAK_GBS,
        "plain": plain,
        "batched_vs_plain": (value / plain["value"]) if plain else None,
        "parity_with_oracle_fixture": parity,
        "cpu_baseline": cpu,
        "device": props["name"],
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
