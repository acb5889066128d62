"""Pins the oracle's three-bit segment step (oracle/or_seg3.c, the restatement
of DB_Opt3BitBFS_Solver::Expand, Common/DiskBasedOptThreeBitBFS.cpp:54-144):

  * a complete DiskBasedOptThreeBitBFS assembled from oracle steps
    (tests/segment_driver.py three_bit_search, following :164-252) reproduces
    the reference's layer counts in the configurations its own tests run:
    sliding 5x2 with 1 and 7 segments (PuzzlesTests/PuzzleSearchTests.cpp:258-333,
    simple and packed index), 9-pancake with segmentBits 15 (:553-560),
    Hanoi 10 with symmetry and segmentBits 20 (:384-390), and the 13-pancake
    with 29-bit segments, with and without the inverse index (:562-590);
    plus smaller segments than the reference uses, so that every one of these
    graphs also runs with many segments.  Hanoi and pancakes have odd cycles,
    so dropping the CurArray exclusion fails them; the 13-pancake's inverse
    index fails without CrossSegmentPostProcess;
  * bucket contents: each (op, segment) bucket decodes to the cross-segment
    children of the new frontier in parent order, recomputed here by brute
    force from the already-pinned single-state expansion;
  * the frontier stream decodes to the written list (the driver checks every
    stream it reads back)."""
import ctypes

import numpy as np
import pytest

import oracle_lib as O
import segment_driver as D

CASES = [
    ("sliding_5x2", "sliding", 22), ("sliding_5x2", "sliding", 19), ("sliding_5x2", "packed", 22),
    ("sliding_5x2", "packed", 19), ("sliding_3x3", "sliding", 14), ("pancake_9", None, 15),
    ("pancake_9", None, 12), ("pancake_5", None, 4), ("hanoi_10_sym", None, 20), ("hanoi_10_sym", None, 12),
    ("hanoi_10", None, 14),
]


def _puzzle(r, variant):
    p = O.from_spec(r["puzzle"])
    if variant == "packed":
        p = O.sliding_packed(p.size, p.height)
    return p


def _nseg(p, sb):
    return max(1, (O.indexes_count(p) + (1 << sb) - 1) >> sb)


@pytest.mark.parametrize("vid,variant,sb", CASES)
def test_three_bit_search_from_segment_steps(oracle, golden, vid, variant, sb):
    r = golden[vid]
    p = _puzzle(r, variant)
    start = O.start_index(p, r["start"])
    counts = D.three_bit_search(p, start, sb, D.oracle_step3(p, sb, _nseg(p, sb)))
    assert counts == r["counts"]


@pytest.mark.parametrize("inv", [False, True])
def test_pancake13_prefix_29bit_segments(oracle, golden, inv):
    r = golden["pancake_13_prefix"]
    p = O.pancake(13, inv_index=inv)
    start = O.start_index(p, r["start"])
    n = len(r["counts"])
    counts = D.three_bit_search(p, start, 29, D.oracle_step3(p, 29, _nseg(p, 29)), max_layers=n)
    assert counts == r["counts"]


def test_inverse_index_needs_postprocess(oracle, golden):
    """A planted mistake: marks of the inverse-index 13-pancake read without
    CrossSegmentPostProcess give other counts (so the test above pins it)."""
    r = golden["pancake_13_prefix"]
    p = O.pancake(13, inv_index=True)
    nseg = _nseg(p, 29)
    q = O.pancake(13, inv_index=False)

    def step(s, cur, old, xs):
        # the marks of the inverse-index puzzle fed to the plain puzzle's step,
        # which applies no post-process
        return O.bfs3_segment_step(q, 29, s, nseg, cur, old, xs) if xs else \
            O.bfs3_segment_step(p, 29, s, nseg, cur, old, xs)
    counts = D.three_bit_search(p, O.start_index(p, r["start"]), 29, step, max_layers=4)
    assert counts != r["counts"][:4]


def test_buckets_are_emission_ordered_chunks(oracle):
    """Buckets of one step hold the cross-segment children of the new frontier
    in parent order, cut into 4096-value chunks."""
    p = O.sliding(5, 2)
    sb = 15
    nseg = _nseg(p, sb)
    rng = np.random.default_rng(3)
    # a frontier of one segment: sorted distinct in-segment ids whose children
    # are mostly new
    seg = 40
    cur = np.unique(rng.integers(0, 1 << sb, size=6000, dtype=np.uint64)).astype(np.uint32)
    r = O.bfs3_segment_step(p, sb, seg, nseg, cur, np.zeros(0, np.uint32))
    assert r["count"] == len(r["fr_idx"]) and np.all(np.diff(r["fr_idx"].astype(np.int64)) > 0)
    expect = {}
    ch = np.zeros(16, np.uint64)
    chop = np.zeros(16, np.int32)
    for i in r["fr_idx"]:
        k = O.lib().or_expand_one(ctypes.byref(p), O.EXPAND_FULL, (seg << sb) | int(i), 0, ch, chop)
        for j in range(k):
            s = int(ch[j]) >> sb
            if s != seg:
                expect.setdefault((int(chop[j]), s), []).append(int(ch[j]) & ((1 << sb) - 1))
    got = {}
    for op, s, nval, data in r["buckets"]:
        assert nval == len(expect[(op, s)])
        got[(op, s)] = D.decode_bucket(data, nval).tolist()
    assert got == expect
    keys = [(op, s) for op, s, _, _ in r["buckets"]]
    assert keys == sorted(keys)
    assert r["children_cross"] == sum(len(v) for v in expect.values())
    # the new frontier: every in-segment child of cur, minus cur itself is not
    # required for a bipartite graph; recompute by brute force
    want = set()
    for i in cur:
        k = O.lib().or_expand_one(ctypes.byref(p), O.EXPAND_FULL, (seg << sb) | int(i), 0, ch, chop)
        for j in range(k):
            if int(ch[j]) >> sb == seg:
                want.add(int(ch[j]) & ((1 << sb) - 1))
    assert sorted(want) == r["fr_idx"].tolist()
    assert np.array_equal(D.decode_fc_stream(r["fr_bytes"]), r["fr_idx"])


def test_no_data_and_old_clear(oracle):
    p = O.pancake(9)
    nseg = _nseg(p, 12)
    r = O.bfs3_segment_step(p, 12, 3, nseg, np.zeros(0, np.uint32), np.array([1, 2, 3], np.uint32))
    assert r["count"] == 0 and len(r["fr_bytes"]) == 0 and not r["buckets"]
    # marks only: the old list removes its ids, odd cycles: cur removes its ids
    xs = [np.array([5, 9, 11], np.uint32)] + [np.zeros(0, np.uint32)] * 7
    r = O.bfs3_segment_step(p, 12, 3, nseg, np.zeros(0, np.uint32), np.array([9], np.uint32), xs)
    assert r["fr_idx"].tolist() == [5, 11]
