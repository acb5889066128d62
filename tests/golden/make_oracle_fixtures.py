"""Runs the CPU oracle once on the full-size workloads and writes their
per-layer counts and frontier hashes to tests/golden/oracle_fixtures.json.

These are the expected values for the GPU parity tests at sizes the CPU suite
cannot afford to recompute.  The counts are pinned where the reference has a
vector (4x3 sliding tile: SlidingPuzzleTests/TestFrontierSearch.cpp:71) and by
closed forms elsewhere (sum = n!, radius = pancake number); the hashes come from
the oracle only.

    python tests/golden/make_oracle_fixtures.py [name ...]
"""
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import oracle_lib as O  # noqa: E402

FIXTURES = {
    # name: (puzzle builder, start text, index_count override, two_bit)
    "pancake_10": (lambda: O.pancake(10), "0 1 2 3 4 5 6 7 8 9", 0, True),
    "pancake_11": (lambda: O.pancake(11), "0 1 2 3 4 5 6 7 8 9 10", 0, False),
    "pancake_12": (lambda: O.pancake(12), "0 1 2 3 4 5 6 7 8 9 10 11", math.factorial(12), False),
    "sliding_4x3": (lambda: O.sliding(4, 3), "0 1 2 3 4 5 6 7 8 9 10 11", 0, False),
    "sliding_3x4": (lambda: O.sliding(3, 4), "0 1 2 3 4 5 6 7 8 9 10 11", 0, False),
    "hanoi_12": (lambda: O.hanoi(12), "index:0", 0, True),
    "hanoi_12_sym": (lambda: O.hanoi(12, True), "index:0", 0, False),
    "hanoi_13": (lambda: O.hanoi(13), "index:0", 0, False),
    # the Hanoi-16 batched search the benchmark times: 4^16 states, 162 layers
    # (about 25 minutes on one core)
    "hanoi_16": (lambda: O.hanoi(16), "index:0", 0, False),
    "sliding_packed_4x3": (lambda: O.sliding_packed(4, 3), "0 1 2 3 4 5 6 7 8 9 10 11", 0, False),
    "sliding_packed_5x2": (lambda: O.sliding_packed(5, 2), "0 1 2 3 4 5 6 7 8 9", 0, True),
}


def main():
    names = sys.argv[1:] or list(FIXTURES)
    dst = os.path.join(HERE, "oracle_fixtures.json")
    data = {}
    if os.path.exists(dst):
        with open(dst) as f:
            data = json.load(f)
    for name in names:
        mk, start_text, nidx, two_bit = FIXTURES[name]
        p = mk()
        start = O.start_index(p, start_text)
        t = time.time()
        counts, hashes = O.bfs3(p, start, index_count=nidx)
        rec = {"puzzle": {"kind": p.kind, "size": p.size, "height": p.height, "flags": p.flags},
               "start": start, "index_count": nidx, "counts": counts, "hashes": [str(h) for h in hashes]}
        if two_bit:
            c2, h2 = O.bfs2(p, start, index_count=nidx)
            assert c2 == counts
            rec["hashes2"] = [str(h) for h in h2]
        rec["oracle_seconds"] = round(time.time() - t, 1)
        data[name] = rec
        print(f"{name}: {len(counts)} layers, sum {sum(counts)}, {rec['oracle_seconds']} s", flush=True)
        with open(dst, "w") as f:
            json.dump(data, f, indent=1)


if __name__ == "__main__":
    main()
