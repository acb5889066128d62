"""GPU parity of the three-bit BFS segment step (ps_seg3_*,
DB_Opt3BitBFS_Solver::Expand, Common/DiskBasedOptThreeBitBFS.cpp:54-144)
against the oracle step, byte for byte: every step of a complete
DiskBasedOptThreeBitBFS assembled from device steps (tests/segment_driver.py)
returns the oracle's layer count, new frontier ids, FrontierCompression
frontier stream and (op, segment) bucket bytes for the same inputs, and the
assembled searches reproduce the reference's own disk-driver vectors
(PuzzlesTests/PuzzleSearchTests.cpp: sliding 5x2 with 1 and 7 segments, simple
and packed index; 9-pancake in 2^15-state segments; Hanoi 10 with symmetry;
the 13-pancake in 29-bit segments with and without the inverse index; the
first layers of Hanoi 21 in 2^32-state segments).  The device decoder reads
every frontier stream back (the compressed input leg of the next layer), and
every bucket stream decodes to the bucket's ids.  Edge cases: a segment with
no data, ids outside the segment, and a workspace reused after a refused step."""
import numpy as np
import pytest

import oracle_lib as O
import puzzlesearch_amd as psa
import segment_driver as D

pytestmark = pytest.mark.gpu

CASES = [
    ("sliding_5x2", "sliding", 22), ("sliding_5x2", "sliding", 19), ("sliding_5x2", "packed", 22),
    ("sliding_5x2", "packed", 19), ("sliding_3x3", "sliding", 14), ("pancake_9", None, 15),
    ("pancake_9", None, 12), ("pancake_5", None, 4), ("hanoi_10_sym", None, 20), ("hanoi_10_sym", None, 12),
    ("hanoi_10", None, 14),
]


@pytest.fixture(scope="module")
def S(ps):
    from puzzlesearch_amd import codecs, seg3
    return seg3, codecs


def _puzzle(r, variant):
    p = O.from_spec(r["puzzle"])
    if variant == "packed":
        p = O.sliding_packed(p.size, p.height)
    return p


def _ps_puzzle(p):
    return psa.Puzzle(p.kind, p.size, p.height, p.flags)


def _nseg(p, sb):
    return max(1, (O.indexes_count(p) + (1 << sb) - 1) >> sb)


def device_step(S, p, sb, check_oracle=True, device_decode=True):
    seg3, codecs = S
    w = seg3.Seg3(_ps_puzzle(p), sb)
    nseg = _nseg(p, sb)
    assert w.nseg == nseg

    def step(s, cur, old, xs):
        r = w.step(s, cur, old, xs)
        if check_oracle:
            o = O.bfs3_segment_step(p, sb, s, nseg, cur, old, xs)
            assert r["count"] == o["count"], (s, r["count"], o["count"])
            assert np.array_equal(r["fr_idx"], o["fr_idx"])
            assert np.array_equal(r["fr_bytes"], o["fr_bytes"])
            assert len(r["buckets"]) == len(o["buckets"])
            for (op, sg, n, b), (oop, osg, on, ob) in zip(r["buckets"], o["buckets"]):
                assert (op, sg, n) == (oop, osg, on)
                assert np.array_equal(b, ob)
        if device_decode and r["count"]:
            chunks = codecs.fc_decode(r["fr_bytes"], r["fr_chunk_offsets"])
            assert np.array_equal(np.concatenate(chunks), r["fr_idx"])
        for (op, sg, n, b), vals in zip(r["buckets"], r["bucket_values"]):
            # the bucket's stream holds its ids in emission order
            assert np.array_equal(D.decode_bucket(b, n), vals)
        return r
    step.workspace = w
    return step


@pytest.mark.parametrize("vid,variant,sb", CASES)
def test_three_bit_search_from_device_steps(S, golden, vid, variant, sb):
    r = golden[vid]
    p = _puzzle(r, variant)
    start = O.start_index(p, r["start"])
    step = device_step(S, p, sb)
    counts = D.three_bit_search(p, start, sb, step)
    step.workspace.close()
    assert counts == r["counts"]


@pytest.mark.parametrize("inv", [False, True])
def test_pancake13_prefix_29bit_segments(S, golden, inv):
    """In-segment and cross-segment flip sets and the inverse-index
    CrossSegmentPostProcess on the device (PuzzleSearchTests.cpp:562-590)."""
    r = golden["pancake_13_prefix"]
    p = O.pancake(13, inv_index=inv)
    step = device_step(S, p, 29)
    counts = D.three_bit_search(p, O.start_index(p, r["start"]), 29, step, max_layers=len(r["counts"]))
    step.workspace.close()
    assert counts == r["counts"]


def test_hanoi21_prefix_32bit_segments(S, golden):
    """1024 segments of 2^32 states (PuzzleSearchTests.cpp:470-477): most
    segments hold no data in these layers and return at once."""
    r = golden["hanoi_21_prefix"]
    p = O.hanoi(21)
    step = device_step(S, p, 32)
    counts = D.three_bit_search(p, 0, 32, step, max_layers=len(r["counts"]))
    step.workspace.close()
    assert counts == r["counts"]


def test_fifteen_puzzle_three_bit_step_full_size(S, ps):
    """The step the benchmark times: the 15-puzzle segment of seed 1234 with
    a Bernoulli(0.2) current frontier and an old frontier of seed 4321 without
    the current ids, against digests of the oracle step's outputs on the same
    inputs (tests/golden/make_segment_fixtures.py --three-bit): count, new
    frontier ids and stream, and every bucket's compressed bytes."""
    import hashlib
    import json
    import os
    from puzzlesearch_amd import workloads as W
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "segment_fixtures.json")
    with open(path) as f:
        rec = json.load(f).get("seg3_fifteen_200000")
    if rec is None:
        pytest.skip("three-bit segment fixture not generated")

    def sha(a):
        return hashlib.sha256(memoryview(np.ascontiguousarray(a)).cast("B")).hexdigest()

    seg3, _ = S
    seg, blank = W.fifteen_segment(rec["seed"])
    cur, _ = W.bernoulli_frontier(rec["seed"], rec["ppm"], blank)
    old0, _ = W.bernoulli_frontier(rec["old_seed"], rec["ppm"], blank)
    old = np.setdiff1d(old0, cur, assume_unique=True).astype(np.uint32)
    assert (seg, len(cur), len(old)) == (rec["segment"], rec["cur"], rec["old"])
    with seg3.Seg3(psa.sliding_packed(4, 4), 32) as w:
        dev = w.step(seg, cur, old)
    assert dev["count"] == rec["count"]
    assert len(dev["fr_idx"]) == rec["frontier"] and sha(dev["fr_idx"]) == rec["fr_idx"]
    assert len(dev["fr_bytes"]) == rec["fr_nbytes"] and sha(dev["fr_bytes"]) == rec["fr_bytes"]
    assert dev["stats"]["children_in"] == rec["children_in"]
    assert dev["stats"]["children_cross"] == rec["children_cross"]
    assert [[o, s, n, len(b), sha(b)] for o, s, n, b in dev["buckets"]] == rec["buckets"]


def test_no_data_out_of_range_and_reuse(S, ps):
    seg3, _ = S
    p = O.pancake(9)
    with seg3.Seg3(psa.pancake(9), 12) as w:
        r = w.step(3, np.zeros(0, np.uint32), np.array([1, 2, 3], np.uint32))
        assert r["count"] == 0 and len(r["fr_bytes"]) == 0 and not r["buckets"]
        with pytest.raises(psa.PSError, match="not below"):
            w.step(3, np.array([5, 1 << 12], np.uint32), np.zeros(0, np.uint32))
        # the refused step left the workspace clean: the next step equals the oracle
        cur = np.array([7, 100, 4000], np.uint32)
        got = w.step(5, cur, np.zeros(0, np.uint32))
        want = O.bfs3_segment_step(p, 12, 5, _nseg(p, 12), cur, np.zeros(0, np.uint32))
        assert np.array_equal(got["fr_idx"], want["fr_idx"])
        assert [b[:3] for b in got["buckets"]] == [b[:3] for b in want["buckets"]]
        with pytest.raises(psa.PSError):
            w.step(_nseg(p, 12), cur, np.zeros(0, np.uint32))


def test_binned_marks_edge_cases(S, ps):
    # the binned path (the 3x3 sliding tile in 2^12-state segments): ids
    # outside the segment among the marks are refused, an old frontier that
    # is not ascending is refused, and after both the workspace still equals
    # the oracle step, incoming marks and old frontier included
    seg3, _ = S
    p = O.sliding(3, 3)
    nseg = _nseg(p, 12)
    with seg3.Seg3(psa.sliding(3, 3), 12) as w:
        xs = [np.zeros(0, np.uint32)] * 3 + [np.array([9, 1 << 12], np.uint32)]
        with pytest.raises(psa.PSError, match="not below"):
            w.step(2, np.array([5], np.uint32), np.zeros(0, np.uint32), xseg=xs)
        with pytest.raises(psa.PSError, match="ascending"):
            w.step(2, np.array([5, 700], np.uint32), np.array([30, 20], np.uint32))
        rng = np.random.default_rng(5)
        cur = np.sort(rng.choice(1 << 12, 300, replace=False)).astype(np.uint32)
        old = np.setdiff1d(np.sort(rng.choice(1 << 12, 200, replace=False)), cur).astype(np.uint32)
        xs = [np.sort(rng.choice(1 << 12, 40, replace=False)).astype(np.uint32) for _ in range(4)]
        got = w.step(2, cur, old, xseg=xs)
        want = O.bfs3_segment_step(p, 12, 2, nseg, cur, old, xs)
        assert got["count"] == want["count"]
        assert np.array_equal(got["fr_idx"], want["fr_idx"])
        assert np.array_equal(got["fr_bytes"], want["fr_bytes"])
        assert [(o, s, n) for o, s, n, _ in got["buckets"]] == [(o, s, n) for o, s, n, _ in want["buckets"]]
        for (_, _, _, a), (_, _, _, b) in zip(got["buckets"], want["buckets"]):
            assert np.array_equal(a, b)
