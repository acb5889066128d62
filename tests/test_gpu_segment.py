"""Device segment step of the frontier search (ps_fs_segment_*) against the
oracle step (oracle/or_segment.c, DB_OptFS_Solver<4>::Expand,
Common/DiskBasedOptFrontierSearch.cpp:53-132), byte for byte: layer count, the
new frontier list and its StreamVInt-with-bounds stream, and every (op,
segment) bucket's FrontierCompression chunks in emission order.

  * complete frontier searches assembled from device steps reproduce the
    reference's layer counts (golden vectors), with every step compared to
    the oracle step on the same inputs, incoming cross-segment marks included
  * the 15-puzzle segment workload at p = 1%, 5%, 20% on frontiers sampled
    across all tile digits, plus one large enough for several 2^18-value
    frontier chunks and 4096-value bucket chunks
  * the workspace is reusable: a second step after a large one, and a step
    after a rejected one, match the oracle"""
import numpy as np
import pytest

import oracle_lib as O
import segment_driver as D

pytestmark = pytest.mark.gpu


def _ps_puzzle(ps, op):
    return ps.Puzzle(op.kind, op.size, op.height, op.flags)


def assert_same(dev, ref, where=""):
    assert dev["count"] == ref["count"], where
    assert np.array_equal(dev["fr_idx"], ref["fr_idx"]), where
    assert np.array_equal(dev["fr_ops"], ref["fr_ops"]), where
    assert np.array_equal(dev["fr_bytes"], ref["fr_bytes"]), where
    assert len(dev["buckets"]) == len(ref["buckets"]), where
    for (o1, s1, n1, b1), (o2, s2, n2, b2) in zip(dev["buckets"], ref["buckets"]):
        assert (o1, s1, n1) == (o2, s2, n2), where
        assert np.array_equal(b1, b2), f"{where} bucket op={o1} seg={s1}"


GOLDEN = [("sliding_3x3", 14), ("sliding_3x2", 6), ("sliding_4x2_edge", 12), ("pancake_5", 6), ("hanoi_10", 16),
          ("hanoi_10_sym", 14)]


@pytest.mark.parametrize("vid,sb", GOLDEN)
def test_frontier_search_golden(ps, golden, vid, sb):
    from puzzlesearch_amd.segment import FsSegment
    r = golden[vid]
    op = O.from_spec(r["puzzle"])
    start = O.start_index(op, r["start"])
    with FsSegment(_ps_puzzle(ps, op), sb) as w:
        nseg = w.nseg

        def step(s, fi, fo, xs):
            dev = w.step(s, fi, fo, xs)
            ref = O.fs_segment_step(op, sb, s, nseg, fi, fo, xs)
            assert_same(dev, ref, f"{vid} segment {s}")
            return dev

        counts = D.frontier_search(op, start, sb, step)
    assert counts == r["counts"]


def fifteen(ppm, stride, seed=1234):
    from puzzlesearch_amd import workloads as W
    seg, b = W.fifteen_segment(seed)
    idx, ops = W.bernoulli_frontier(seed, ppm, b, m_stride=stride)
    return seg, idx, ops


@pytest.mark.parametrize("ppm", [10000, 50000, 200000])
def test_fifteen_puzzle_segment(ps, ppm):
    from puzzlesearch_amd.segment import FsSegment
    op = O.sliding_packed(4, 4)
    seg, idx, ops = fifteen(ppm, 4099)
    with FsSegment(_ps_puzzle(ps, op), 32) as w:
        assert w.nseg == 3805
        dev = w.step(seg, idx, ops)
    ref = O.fs_segment_step(op, 32, seg, 3805, idx, ops)
    assert_same(dev, ref, f"ppm={ppm}")
    st = dev["stats"]
    assert st["children_in"] == ref["children_in"] and st["children_cross"] == ref["children_cross"]


def _segment_fixtures():
    import json
    import os
    p = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "segment_fixtures.json")
    if not os.path.exists(p):
        return {}
    with open(p) as f:
        return json.load(f)


SEG_FIX = _segment_fixtures()


@pytest.mark.parametrize("ppm", [10000, 50000, 200000])
def test_fifteen_puzzle_segment_full_size(ps, ppm):
    """The step the benchmark times -- seed 1234, every id of the segment's
    blank cell (stride 1) -- against digests of the oracle step's outputs on
    the same frontier (tests/golden/make_segment_fixtures.py): count, new
    frontier ids, op bits and stream, and every bucket's compressed bytes."""
    import hashlib
    from puzzlesearch_amd.segment import FsSegment
    key = f"fifteen_{ppm}"
    if key not in SEG_FIX:
        pytest.skip("segment fixture not generated")
    rec = SEG_FIX[key]

    def sha(a):
        return hashlib.sha256(memoryview(np.ascontiguousarray(a)).cast("B")).hexdigest()

    op = O.sliding_packed(4, 4)
    seg, idx, ops = fifteen(ppm, 1, seed=rec["seed"])
    assert (seg, len(idx)) == (rec["segment"], rec["parents"])
    with FsSegment(_ps_puzzle(ps, op), 32) as w:
        dev = w.step(seg, idx, ops)
    assert dev["count"] == rec["count"]
    assert len(dev["fr_idx"]) == rec["frontier"]
    assert sha(dev["fr_idx"]) == rec["fr_idx"]
    assert sha(dev["fr_ops"]) == rec["fr_ops"]
    assert len(dev["fr_bytes"]) == rec["fr_nbytes"] and sha(dev["fr_bytes"]) == rec["fr_bytes"]
    assert dev["stats"]["children_in"] == rec["children_in"]
    assert dev["stats"]["children_cross"] == rec["children_cross"]
    assert [[o, s, n, len(b), sha(b)] for o, s, n, b in dev["buckets"]] == rec["buckets"]


def test_fifteen_puzzle_every_blank_cell(ps):
    # parents with the blank in every cell, so each of the cross pass's slot
    # skips (horizontal moves, vertical moves left of list position 12) is
    # checked against the oracle's full expansion
    from puzzlesearch_amd.segment import FsSegment
    op = O.sliding_packed(4, 4)
    rng = np.random.default_rng(7)
    seg = 0x0A5  # t12 = 5, t13 = 10, t14 = 0
    m = np.unique(rng.integers(0, 3832012800 // 16, size=60000, dtype=np.uint64))
    idx = (m * np.uint64(16) + rng.integers(0, 16, size=m.size, dtype=np.uint64)).astype(np.uint32)
    ops = rng.integers(0, 16, size=idx.size, dtype=np.uint64).astype(np.uint8)
    assert set((idx & 15).tolist()) == set(range(16))
    with FsSegment(_ps_puzzle(ps, op), 32) as w:
        dev = w.step(seg, idx, ops)
    ref = O.fs_segment_step(op, 32, seg, 3805, idx, ops)
    assert_same(dev, ref, "every blank cell")
    assert dev["stats"]["children_cross"] == ref["children_cross"] > 0


def test_fifteen_puzzle_chunking_and_reuse(ps):
    from puzzlesearch_amd.segment import FsSegment
    op = O.sliding_packed(4, 4)
    seg, idx, ops = fifteen(200000, 37)
    small_seg, sidx, sops = fifteen(50000, 4099, seed=99)
    with FsSegment(_ps_puzzle(ps, op), 32) as w:
        dev = w.step(seg, idx, ops)
        assert dev["stats"]["frontier_chunks"] > 1
        assert dev["stats"]["cross_chunks"] > dev["stats"]["buckets"]
        ref = O.fs_segment_step(op, 32, seg, 3805, idx, ops)
        assert_same(dev, ref, "large")
        # incoming cross-segment marks on a second, smaller step
        rng = np.random.default_rng(3)
        xseg = [np.sort(rng.integers(0, 3832012800, size=k, dtype=np.uint64)).astype(np.uint32)
                for k in (100, 0, 2000, 37)]
        dev2 = w.step(small_seg, sidx, sops, xseg)
        ref2 = O.fs_segment_step(op, 32, small_seg, 3805, sidx, sops, xseg)
        assert_same(dev2, ref2, "second step")


def test_rejected_ids_leave_workspace_clean(ps):
    from puzzlesearch_amd.segment import FsSegment
    op = O.sliding(3, 3)
    with FsSegment(_ps_puzzle(ps, op), 12) as w:
        bad = np.array([5, 4096 + 7], np.uint32)  # 4103 is not below 2^12
        with pytest.raises(ps.PSError):
            w.step(0, bad, np.zeros(2, np.uint8))
        good = np.array([5, 100, 2000], np.uint32)
        dev = w.step(3, good, np.zeros(3, np.uint8))
        ref = O.fs_segment_step(op, 12, 3, w.nseg, good, np.zeros(3, np.uint8))
        assert_same(dev, ref, "after rejection")


def test_too_many_moves_is_unsupported(ps):
    from puzzlesearch_amd.segment import FsSegment
    with pytest.raises(ps.PSError):
        FsSegment(ps.pancake(9), 16)
