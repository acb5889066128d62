"""Serving loop over the multi-search engine: a seeded Poisson stream of
complete-BFS requests drawn from the serving mix, admitted by HBM footprint
and batched on the device while they run (puzzlesearch_amd.Batch).

    res = serve(rate=20, n_jobs=40, seed=7)
    res["p50_ms"], res["p99_ms"], res["states_per_s"]

A request's latency is its completion time minus its arrival time.  The loop
submits every request whose arrival time has passed, steps the engine while
anything is queued or running, and sleeps only when the device is idle."""
import math
import time

import numpy as np

from . import Batch, JOB_DONE, hanoi, pancake, sliding

# sliding tile 2x4, 2x5, 3x3, 3x4; pancake 8..12; 4-peg Hanoi 10..16 disks (16 kinds)
MIX = ([("sliding", 2, 4), ("sliding", 2, 5), ("sliding", 3, 3), ("sliding", 3, 4)]
       + [("pancake", n, 0) for n in range(8, 13)]
       + [("hanoi", d, 0) for d in range(10, 17)])


def make_puzzle(kind, a, b):
    if kind == "sliding":
        return sliding(a, b), math.factorial(a * b)
    if kind == "pancake":
        return pancake(a), math.factorial(a)
    return hanoi(a), 4 ** a


def reachable(kind, a, b):
    """States a complete search from the solved state visits."""
    if kind == "sliding":
        return math.factorial(a * b) // 2
    if kind == "pancake":
        return math.factorial(a)
    return 4 ** a


def draw(n_jobs, rate, seed):
    """(arrival seconds, mix entry) for each request: exponential gaps with mean
    1/rate and kinds uniform over MIX, from one seeded generator."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=n_jobs)
    kinds = rng.integers(0, len(MIX), size=n_jobs)
    return [(float(t), MIX[int(k)]) for t, k in zip(np.cumsum(gaps), kinds)]


def serve(rate, n_jobs, seed, max_jobs=64, segment_bits=22, hbm_budget=0, requests=None, job_grid_blocks=None,
          reserved_cus=None, shortest_first=None):
    reqs = requests if requests is not None else draw(n_jobs, rate, seed)
    with Batch(max_jobs=max_jobs, segment_bits=segment_bits, hbm_budget=hbm_budget,
               job_grid_blocks=job_grid_blocks, reserved_cus=reserved_cus, shortest_first=shortest_first) as b:
        t0 = time.perf_counter()
        pending = list(enumerate(reqs))
        arrival, ids = {}, {}
        done_at = {}
        max_running = 0
        while pending or len(done_at) < len(reqs):
            now = time.perf_counter() - t0
            while pending and pending[0][1][0] <= now:
                i, (t, (kind, a, c)) = pending.pop(0)
                p, n = make_puzzle(kind, a, c)
                jid = b.submit(p, 0, n)
                ids[jid] = i
                arrival[i] = t
            running, finished, queued = b.step()
            max_running = max(max_running, running)
            t_step = time.perf_counter() - t0
            for jid in b.finished():
                done_at[ids[jid]] = (t_step, jid)
            if running == 0 and queued == 0 and pending:
                time.sleep(max(0.0, min(0.01, pending[0][1][0] - (time.perf_counter() - t0))))
        makespan = time.perf_counter() - t0
        results = []
        for i, (t, (kind, a, c)) in enumerate(reqs):
            t_done, jid = done_at[i]
            info = b.job(jid)
            results.append({"request": i, "kind": f"{kind}({a},{c})" if kind == "sliding" else f"{kind}({a})",
                            "arrival_s": t, "latency_ms": (t_done - arrival[i]) * 1e3, "state": info["state"],
                            "counts": info["counts"], "hashes": info["hashes"], "steps": info["steps"],
                            "expected_states": reachable(kind, a, c)})
        stats = b.stats()
    lat = np.array([r["latency_ms"] for r in results])
    ok = all(r["state"] == JOB_DONE and sum(r["counts"]) == r["expected_states"] for r in results)
    return {"rate": rate, "jobs": len(reqs), "seed": seed, "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()), "makespan_s": makespan,
            "states": int(sum(r["expected_states"] for r in results)),
            "states_per_s": sum(r["expected_states"] for r in results) / makespan,
            "max_running": max_running, "engine": stats, "all_complete_and_summed": ok, "results": results}
