"""Autotuner for the device BFS: picks the launch and geometry options of
ps_bfs for one puzzle on the current device by measurement.

Searched options (ps_bfs_options):
  block_threads   threads per expand block: 256, 512
  tile_bits_log2  LDS window: 0 = the puzzle's locality block, k = 2^k states, -1 = none
  grid_blocks     persistent expand blocks: 0 = resident blocks per CU x CUs, or an explicit count
  segment_bits    segment size 2^s (0 = one segment)
  cross_marks     PS_CROSS_BYTES / PS_CROSS_ATOMIC
  skip_visited    0 / 1
  graph_layers    layers captured per hipGraph replay
  direction       0 = top-down layers only, 1 = bottom-up when the frontier outnumbers the unvisited states

The search is coordinate descent from the defaults: one option at a time over
its candidate values, keeping the fastest, for `passes` rounds.  Every
candidate's layer counts and frontier hashes must equal the first run's: a
configuration that changes the result raises instead of being picked.  Results are
stored as JSON keyed by device name and puzzle, e.g.

    {"gfx950 ...": {"pancake(12)": {"options": {...}, "ms": 17.3, "tried": 23}}}
"""
import json
import os
import time

from . import Bfs, Puzzle, sync

CANDIDATES = {
    "block_threads": [512, 256],
    "tile_bits_log2": [0, -1, 12, 14, 16],
    "grid_blocks": [0, 256, 512, 1024],
    "segment_bits": [22, 0, 20, 24, 26],
    "cross_marks": [0, 1],
    "skip_visited": [0, 1],
    "graph_layers": [4, 1, 8, 16],
    "direction": [0, 1],
    "fused_layers": [0, 1],
}
DEFAULTS = {k: v[0] for k, v in CANDIDATES.items()}


def puzzle_key(p):
    return repr(p)


def measure(p, start, opts, runs=2, index_count=0):
    """Best-of-`runs` ms per complete search, with the counts and hashes."""
    kw = dict(opts)
    if index_count:
        kw["index_count"] = index_count
    with Bfs(p, graph=1, **kw) as b:
        counts, hashes = b.run(start)  # warm-up, graph capture
        best = None
        for _ in range(runs):
            sync()
            t0 = time.perf_counter()
            c2, h2 = b.run(start)
            sync()
            dt = (time.perf_counter() - t0) * 1e3
            if (c2, h2) != (counts, hashes):
                raise RuntimeError(f"non-deterministic result for {opts}")
            best = dt if best is None else min(best, dt)
    return best, counts, hashes


def tune(p, start, index_count=0, passes=1, runs=2, candidates=None, log=None):
    cand = dict(CANDIDATES)
    if candidates:
        cand.update(candidates)
    cur = {k: v[0] for k, v in cand.items()}
    best_ms, ref_counts, ref_hashes = measure(p, start, cur, runs, index_count)
    tried = {json.dumps(cur, sort_keys=True): best_ms}
    if log:
        log(f"tune {p}: start {cur} -> {best_ms:.3f} ms")
    for _ in range(passes):
        for name, values in cand.items():
            for v in values:
                trial = dict(cur)
                trial[name] = v
                key = json.dumps(trial, sort_keys=True)
                if key in tried:
                    continue
                try:
                    ms, counts, hashes = measure(p, start, trial, runs, index_count)
                except Exception as e:  # an unsupported combination (e.g. LDS too large)
                    tried[key] = None
                    if log:
                        log(f"  {name}={v}: rejected ({e})")
                    continue
                if (counts, hashes) != (ref_counts, ref_hashes):
                    raise RuntimeError(f"options {trial} changed the search result")
                tried[key] = ms
                if log:
                    log(f"  {name}={v}: {ms:.3f} ms")
                if ms < best_ms:
                    best_ms, cur = ms, trial
    return {"options": cur, "ms": best_ms, "tried": len(tried), "counts": ref_counts}


def cache_path():
    return os.environ.get("PS_TUNING_FILE",
                          os.path.join(os.path.expanduser("~"), ".cache", "puzzlesearch_amd", "tuning.json"))


def load(device, p, path=None):
    path = path or cache_path()
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get(device, {}).get(puzzle_key(p))


def store(device, p, result, path=None):
    path = path or cache_path()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    data = {}
    if os.path.exists(path):
        with open(path) as f:
            data = json.load(f)
    entry = {k: v for k, v in result.items() if k != "counts"}
    data.setdefault(device, {})[puzzle_key(p)] = entry
    with open(path, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    return path


def tuned_options(device, p, start, index_count=0, path=None, log=None, **kw):
    """The stored options for this device and puzzle, tuning (and storing) them first if absent."""
    hit = load(device, p, path)
    if hit:
        return hit["options"], hit
    res = tune(p, start, index_count=index_count, log=log, **kw)
    store(device, p, res, path)
    return res["options"], res


def as_puzzle(kind, size, height=0, flags=0):
    return Puzzle(kind, size, height, flags)
