"""CPU oracle pinned against the reference's golden layer-count vectors
(tests/golden/reference_vectors.json, each entry records its source file:line),
closed forms and brute force.

What pins what:
  * or_bfs3 / or_bfs2 (InMemoryThreeBitBFS / InMemoryTwoBitBFS restatements):
    every full golden vector small enough for the CPU suite, from both drivers.
  * or_bfs3_segmented (DiskBasedOptThreeBitBFS restatement): the reference
    tests' own multi-segment runs (PuzzlesTests/PuzzleSearchTests.cpp).
  * partial vectors (21-disk Hanoi, 13-pancake prefixes): a plain hash-set BFS
    over the oracle's move generator -- a different algorithm than the bit-array
    drivers, so it also cross-checks them.
  * closed forms: sum of layers = reachable states (n!, 4^n, (wh)!/2), pancake
    radius = pancake number P(n), sliding-tile layers bipartite.
  * rank/unrank: lexicographic position (itertools order) for every
    permutation of n <= 8, and the reference's own PermutationMap.cpp built into
    oracle/_ref when the reference tree is present.
"""
import ctypes
import itertools
import math
import os

import numpy as np
import pytest

import oracle_lib as O

# pancake numbers P(n) (maximum flips needed), OEIS A058986
PANCAKE_NUMBER = {2: 1, 3: 3, 4: 4, 5: 5, 6: 7, 7: 8, 8: 9, 9: 10, 10: 11, 11: 13, 12: 14}

CPU_SIZED = ["sliding_2x2", "sliding_3x2", "sliding_3x2_edge", "sliding_4x2", "sliding_4x2_edge",
             "sliding_3x3", "sliding_3x3_edge", "sliding_3x3_center", "sliding_5x2",
             "pancake_5", "pancake_9", "hanoi_10", "hanoi_10_sym"]


@pytest.mark.parametrize("vid", CPU_SIZED)
def test_bfs3_matches_golden(oracle, golden, vid):
    r = golden[vid]
    p = O.from_spec(r["puzzle"])
    counts, hashes = O.bfs3(p, O.start_index(p, r["start"]))
    assert counts == r["counts"], r["source"]
    assert len(set(hashes)) == len(hashes)  # distinct frontiers hash differently


@pytest.mark.parametrize("vid", CPU_SIZED)
def test_bfs2_matches_golden(oracle, golden, vid):
    r = golden[vid]
    p = O.from_spec(r["puzzle"])
    counts, _ = O.bfs2(p, O.start_index(p, r["start"]))
    assert counts == r["counts"], r["source"]


@pytest.mark.parametrize("vid,seg_bits", [("pancake_9", 15), ("pancake_9", 32), ("hanoi_10_sym", 20),
                                          ("hanoi_10", 17), ("sliding_5x2", 19)])
def test_segmented_matches_golden(oracle, golden, vid, seg_bits):
    r = golden[vid]
    p = O.from_spec(r["puzzle"])
    assert O.bfs3_segmented(p, O.start_index(p, r["start"]), seg_bits) == r["counts"], r["source"]


@pytest.mark.parametrize("inv", [False, True])
def test_segmented_pancake13_prefix(oracle, golden, inv):
    # PuzzlesTests/PuzzleSearchTests.cpp: 29-bit segments, maxSteps 6, with and
    # without the inverse index (cross-segment post-process)
    r = golden["pancake_13_prefix"]
    p = O.pancake(13, inv)
    assert O.bfs3_segmented(p, 0, 29, max_layers=len(r["counts"])) == r["counts"]


def sparse_bfs(p, start, layers):
    """Plain set-based BFS over the oracle's move generator."""
    seen_prev, cur = set(), {start}
    out = [1]
    visited = {start}
    for _ in range(layers - 1):
        nxt = set()
        for s in cur:
            ch, _ = O.expand_one(p, s)
            for c in ch:
                if c not in visited:
                    nxt.add(c)
        visited |= nxt
        out.append(len(nxt))
        cur = nxt
    return out


@pytest.mark.parametrize("vid", ["hanoi_21_prefix", "hanoi_21_sym_prefix"])
def test_sparse_prefix_hanoi21(oracle, golden, vid):
    r = golden[vid]
    p = O.from_spec(r["puzzle"])
    assert sparse_bfs(p, 0, len(r["counts"])) == r["counts"], r["source"]


@pytest.mark.parametrize("vid", ["hanoi_17_120", "hanoi_18_120", "hanoi_19_120"])
def test_sparse_prefix_hanoi_vertical(oracle, golden, vid):
    # the first 24 of the 121 layers of HanoiTowersTests/HanoiSearchTests.cpp
    r = golden[vid]
    p = O.from_spec(r["puzzle"])
    assert sparse_bfs(p, 0, 24) == r["counts"][:24], r["source"]


def test_sparse_prefix_pancake13(oracle, golden):
    r = golden["pancake_13_prefix"]
    assert sparse_bfs(O.pancake(13), 0, 5) == r["counts"][:5]


@pytest.mark.parametrize("vid", ["pancake_9", "hanoi_10_sym", "sliding_3x3"])
def test_sparse_equals_bitarray_driver(oracle, golden, vid):
    r = golden[vid]
    p = O.from_spec(r["puzzle"])
    s = O.start_index(p, r["start"])
    counts, _ = O.bfs3(p, s)
    assert sparse_bfs(p, s, len(counts)) == counts


def test_closed_forms(oracle, golden):
    for vid, r in golden.items():
        if r["partial"]:
            continue
        pz = r["puzzle"]
        total = sum(r["counts"])
        if pz["kind"] == "pancake":
            assert total == math.factorial(pz["n"])
            assert len(r["counts"]) - 1 == PANCAKE_NUMBER[pz["n"]]
        elif pz["kind"] == "sliding":
            assert total == math.factorial(pz["width"] * pz["height"]) // 2
        elif pz["kind"] == "hanoi" and not pz["symmetry"]:
            assert total == 4 ** pz["disks"]


@pytest.mark.parametrize("n", [4, 5, 6, 7, 8])
def test_pancake_small_full(oracle, n):
    counts, _ = O.bfs3(O.pancake(n), 0)
    assert sum(counts) == math.factorial(n)
    assert len(counts) - 1 == PANCAKE_NUMBER[n]
    assert counts[1] == n - 1 and counts[2] == (n - 1) * (n - 2)


def test_sliding_bipartite(oracle):
    # a sliding-tile move changes the blank's row+column parity, so every
    # child of layer d is in layer d+1 or d-1: check via the frontier of 3x3
    p = O.sliding(3, 3)
    s = O.parse(p, "0 1 2 3 4 5 6 7 8")
    counts, _ = O.bfs3(p, s)
    counts2, _ = O.bfs2(p, s)
    assert counts == counts2


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8])
def test_rank_is_lexicographic_position(oracle, n):
    L = O.lib()
    for pos, perm in enumerate(itertools.permutations(range(n))):
        arr = np.array(perm, dtype=np.int32)
        assert L.or_perm_rank(arr.copy(), n) == pos
        out = np.zeros(n, dtype=np.int32)
        L.or_perm_unrank(pos, out, n)
        assert tuple(out) == perm


def test_rank_large_random(oracle):
    L = O.lib()
    rng = np.random.default_rng(7)
    for n in (12, 13, 16):
        for _ in range(200):
            perm = rng.permutation(n).astype(np.int32)
            idx = L.or_perm_rank(perm.copy(), n)
            # lexicographic rank by definition: sum of (#smaller to the right) * (n-1-i)!
            expect = sum(sum(1 for y in perm[i + 1:] if y < perm[i]) * math.factorial(n - 1 - i) for i in range(n))
            assert idx == expect
            out = np.zeros(n, dtype=np.int32)
            L.or_perm_unrank(idx, out, n)
            assert (out == perm).all()


REF_SO = os.path.join(O.ORACLE_DIR, "_ref", "libpermref.so")


@pytest.mark.skipif(not os.path.exists(REF_SO), reason="oracle/_ref not built (reference tree absent)")
def test_rank_against_reference_build(oracle):
    """Puzzles/PermutationMap.cpp itself, compiled from the reference tree."""
    R = ctypes.CDLL(REF_SO)
    rank = R._Z15PermutationRankPii
    rank.restype = ctypes.c_uint64
    rank.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    unrank = R._Z17PermutationUnrankmPii
    unrank.restype = None
    unrank.argtypes = [ctypes.c_uint64, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    L = O.lib()
    rng = np.random.default_rng(3)
    for n in (3, 6, 9, 12, 16):
        for _ in range(300):
            perm = rng.permutation(n).astype(np.int32)
            a = (ctypes.c_int * n)(*perm.tolist())
            assert rank(a, n) == L.or_perm_rank(perm.copy(), n)
            idx = rng.integers(0, math.factorial(n), dtype=np.uint64) if n < 20 else 0
            b = (ctypes.c_int * 16)()
            unrank(int(idx), b, n)
            out = np.zeros(16, dtype=np.int32)
            L.or_perm_unrank(int(idx), out, n)
            assert list(b)[:n] == out[:n].tolist()
