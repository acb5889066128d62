"""GPU parity of the device-resident BFS (ps_bfs_*) with the reference
drivers: per-layer counts must equal the reference's golden vectors, and
per-layer frontier hashes must equal the oracle's, in every execution mode:
batched (persistent kernels, with and without the captured hipGraph), plain
(one launch per segment per layer), several segment sizes, LDS window on and
off, and the two-bit driver.  Full-size workloads are compared with the oracle
fixtures in tests/golden/oracle_fixtures.json."""
import json
import math
import os

import numpy as np
import pytest

import oracle_lib as O
import puzzlesearch_amd as psa

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))

SMALL = ["sliding_3x2", "sliding_4x2_edge", "sliding_3x3", "sliding_3x3_center", "sliding_5x2",
         "pancake_5", "pancake_9", "hanoi_10", "hanoi_10_sym"]

MODES = {
    "batched_graph": dict(mode=psa.BFS_BATCHED, graph=1),
    "batched_eager": dict(mode=psa.BFS_BATCHED, graph=0),
    "plain": dict(mode=psa.BFS_PLAIN),
    "batched_seg12": dict(mode=psa.BFS_BATCHED, segment_bits=12),
    # tiles of 2^19 states over 2^12-state segments (128 segments a tile): the
    # bookkeeping kernel lists such tiles without reading every segment
    "batched_tile19_seg12": dict(mode=psa.BFS_BATCHED, segment_bits=12, tile_bits_log2=19),
    "plain_seg12": dict(mode=psa.BFS_PLAIN, segment_bits=12),
    "batched_nowindow": dict(mode=psa.BFS_BATCHED, tile_bits_log2=-1),
    "batched_tile12_256": dict(mode=psa.BFS_BATCHED, tile_bits_log2=12, block_threads=256),
    "batched_atomic": dict(mode=psa.BFS_BATCHED, cross_marks=psa.CROSS_ATOMIC),
    "plain_seg12_atomic": dict(mode=psa.BFS_PLAIN, segment_bits=12, cross_marks=psa.CROSS_ATOMIC),
    "batched_auto_dir": dict(mode=psa.BFS_BATCHED, direction=psa.DIR_AUTO),
    "batched_bottom_up": dict(mode=psa.BFS_BATCHED, direction=psa.DIR_BOTTOM_UP),
    "batched_bottom_up_seg12": dict(mode=psa.BFS_BATCHED, direction=psa.DIR_BOTTOM_UP, segment_bits=12),
    # finalize and expansion of each tile in one launch (four rotating arrays):
    # one segment; one tile per segment; four tiles per segment, eager
    "batched_fused": dict(mode=psa.BFS_BATCHED, cross_marks=psa.CROSS_ATOMIC, fused_layers=1),
    "batched_fused_tile12_seg12": dict(mode=psa.BFS_BATCHED, cross_marks=psa.CROSS_ATOMIC, fused_layers=1,
                                       segment_bits=12, tile_bits_log2=12),
    "batched_fused_tile12_seg14_eager": dict(mode=psa.BFS_BATCHED, cross_marks=psa.CROSS_ATOMIC, fused_layers=1,
                                             segment_bits=14, tile_bits_log2=12, graph=0, block_threads=256),
}


def _ps_puzzle(op):
    return psa.Puzzle(op.kind, op.size, op.height, op.flags)


_oracle_cache = {}


def oracle_run(vid, golden, bits=3):
    key = (vid, bits)
    if key not in _oracle_cache:
        r = golden[vid]
        op = O.from_spec(r["puzzle"])
        s = O.start_index(op, r["start"])
        _oracle_cache[key] = (O.bfs3 if bits == 3 else O.bfs2)(op, s)
    return _oracle_cache[key]


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("vid", SMALL)
def test_bfs3_golden(ps, golden, vid, mode):
    r = golden[vid]
    op = O.from_spec(r["puzzle"])
    start = O.start_index(op, r["start"])
    if MODES[mode].get("direction") == psa.DIR_BOTTOM_UP and vid == "hanoi_10_sym":
        # non-canonical indexes are not states of the reduced graph
        with pytest.raises(psa.PSError, match="bottom-up"):
            psa.Bfs(_ps_puzzle(op), **MODES[mode])
        return
    with psa.Bfs(_ps_puzzle(op), **MODES[mode]) as b:
        counts, hashes = b.run(start)
        st = b.stats()
    assert counts == r["counts"], r["source"]
    _, ohash = oracle_run(vid, golden)
    assert hashes == ohash
    if MODES[mode].get("direction", 0) == 0:
        assert st.expanded == sum(r["counts"])  # top-down: every state expanded once
    else:
        assert st.expanded <= sum(r["counts"])  # bottom-up layers expand no parents


@pytest.mark.parametrize("vid", SMALL)
def test_bfs2_golden(ps, golden, vid):
    r = golden[vid]
    op = O.from_spec(r["puzzle"])
    start = O.start_index(op, r["start"])
    with psa.Bfs(_ps_puzzle(op), bits=2) as b:
        counts, hashes = b.run(start)
    assert counts == r["counts"], r["source"]
    _, ohash = oracle_run(vid, golden, bits=2)
    assert hashes == ohash


def test_graph_replays_repeat(ps, golden):
    """The captured layer graph gives identical results on every replay, with
    eager runs and synchronisations in between."""
    r = golden["pancake_9"]
    p = psa.pancake(9)
    with psa.Bfs(p, graph=1) as g, psa.Bfs(p, graph=0) as e:
        ref = e.run(0)
        for _ in range(4):
            assert g.run(0) == ref
            assert e.run(0) == ref
            psa.sync()
    assert ref[0] == r["counts"]


def test_step_api_matches_run(ps, golden):
    r = golden["hanoi_10_sym"]
    with psa.Bfs(psa.hanoi(10, True)) as b:
        b.reset(0)
        counts = [1]
        while True:
            c, _ = b.step()
            if c == 0:
                break
            counts.append(c)
    assert counts == r["counts"]


def test_frontier_array_matches_oracle(ps):
    """Final frontier bit array equals the oracle's last layer."""
    p = psa.pancake(8)
    with psa.Bfs(p, graph=0) as b:
        counts, hashes = b.run(0)
        words = b.frontier()
    assert O.bits_hash(words) == hashes[-1]
    assert int(np.unpackbits(words.view(np.uint8)).sum()) == counts[-1]


def test_max_layers(ps, golden):
    r = golden["pancake_9"]
    with psa.Bfs(psa.pancake(9), max_layers=4) as b:
        counts, _ = b.run(0)
    assert counts == r["counts"][:4]


def test_start_out_of_range(ps):
    with psa.Bfs(psa.pancake(7)) as b:
        with pytest.raises(psa.PSError) as e:
            b.run(math.factorial(7))
        assert e.value.code == psa.PS_ERR_INDEX_RANGE


def test_index_count_too_small_is_reported(ps):
    # an index space smaller than the puzzle's: children beyond it are refused
    with psa.Bfs(psa.pancake(8), index_count=1000) as b:
        with pytest.raises(psa.PSError) as e:
            b.run(0)
        assert e.value.code == psa.PS_ERR_INDEX_RANGE


def _fixtures():
    path = os.path.join(HERE, "golden", "oracle_fixtures.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


FIX = _fixtures()


@pytest.mark.parametrize("name", [n for n in ("pancake_10", "pancake_11", "hanoi_12", "hanoi_12_sym", "hanoi_13",
                                                "sliding_packed_5x2", "sliding_packed_4x3") if n in FIX])
@pytest.mark.parametrize("mode", ["batched_graph", "plain_seg20", "fused_seg20"])
def test_medium_against_oracle_fixture(ps, name, mode):
    rec = FIX[name]
    p = psa.Puzzle(**rec["puzzle"])
    opts = dict(MODES.get(mode, {}))
    if mode == "plain_seg20":
        opts = dict(mode=psa.BFS_PLAIN, segment_bits=20)
    if mode == "fused_seg20":
        # tiles of 2^16 states inside 2^20-state segments
        opts = dict(cross_marks=psa.CROSS_ATOMIC, fused_layers=1, segment_bits=20, tile_bits_log2=16)
    with psa.Bfs(p, index_count=rec["index_count"], **opts) as b:
        counts, hashes = b.run(rec["start"])
    assert counts == rec["counts"]
    assert [str(h) for h in hashes] == rec["hashes"]


@pytest.mark.parametrize("name", [n for n in ("pancake_10", "hanoi_12", "sliding_packed_5x2") if n in FIX])
def test_two_bit_against_oracle_fixture(ps, name):
    rec = FIX[name]
    p = psa.Puzzle(**rec["puzzle"])
    with psa.Bfs(p, bits=2, index_count=rec["index_count"]) as b:
        counts, hashes = b.run(rec["start"])
    assert counts == rec["counts"]
    assert [str(h) for h in hashes] == rec["hashes2"]


@pytest.mark.parametrize("seg,cross,direction", [(0, 0, 0), (20, 0, 0), (26, 0, 0), (22, 1, 0), (22, 0, 1), (20, 0, 2)])
def test_pancake12_full(ps, seg, cross, direction):
    """12-pancake, 12! states, three-bit, identity start: counts and frontier
    hashes equal the oracle; sum = 12!, radius = P(12) = 14; byte and atomic
    cross-window marks."""
    if "pancake_12" not in FIX:
        pytest.skip("oracle fixture pancake_12 not generated")
    rec = FIX["pancake_12"]
    with psa.Bfs(psa.pancake(12), index_count=math.factorial(12), segment_bits=seg, cross_marks=cross,
                 direction=direction) as b:
        counts, hashes = b.run(0)
    assert sum(counts) == math.factorial(12)
    assert len(counts) == 15
    assert counts == rec["counts"]
    assert [str(h) for h in hashes] == rec["hashes"]


@pytest.mark.parametrize("inv", [False, True])
def test_pancake13_prefix(ps, golden, inv):
    """13-pancake in PancakeOptimized's 29-bit segmented index space
    (13 x 2^29 indexes): the first layers equal the reference's prefix
    (PuzzlesTests/PuzzleSearchTests.cpp:569) and the oracle's hashes."""
    r = golden["pancake_13_prefix"]
    n_layers = len(r["counts"])
    op = O.pancake(13, inv_index=inv)
    start = O.start_index(op, r["start"])
    with psa.Bfs(psa.pancake(13, inv_index=inv), segment_bits=29, max_layers=n_layers) as b:
        counts, hashes = b.run(start)
    assert counts == r["counts"]
    oc, oh = O.bfs3(op, start, max_layers=n_layers)
    assert (counts, hashes) == (oc, oh)


def test_bottom_up_index_space_past_last_state(ps, golden):
    """An index space longer than the state space (9-pancake with
    index_count = 9! + 64, or the 12-pancake's 2^29 IndexesCount()) holds
    indexes that are no state: bottom-up layers are refused there, and AUTO
    stays top-down and equals the oracle."""
    r = golden["pancake_9"]
    n = math.factorial(9) + 64
    with pytest.raises(psa.PSError, match="bottom-up"):
        psa.Bfs(psa.pancake(9), index_count=n, direction=psa.DIR_BOTTOM_UP)
    with pytest.raises(psa.PSError, match="bottom-up"):
        psa.Bfs(psa.pancake(12), direction=psa.DIR_BOTTOM_UP)
    with psa.Bfs(psa.pancake(9), index_count=n, direction=psa.DIR_AUTO) as b:
        counts, hashes = b.run(0)
    oc, oh = O.bfs3(O.pancake(9), 0, index_count=n)
    assert counts == r["counts"] == oc
    assert hashes == oh


def test_pancake12_reference_sizing(ps):
    """Same search with PancakeOptimized's 2^29 index space: identical result."""
    if "pancake_12" not in FIX:
        pytest.skip("oracle fixture pancake_12 not generated")
    rec = FIX["pancake_12"]
    with psa.Bfs(psa.pancake(12), mode=psa.BFS_PLAIN, segment_bits=24) as b:
        counts, hashes = b.run(0)
    assert counts == rec["counts"]
    assert [str(h) for h in hashes] == rec["hashes"]


@pytest.mark.parametrize("shape,cross", [("sliding_4x3", 0), ("sliding_3x4", 0), ("sliding_4x3", 1)])
def test_sliding_4x3_full(ps, golden, shape, cross):
    """4x3 sliding tile, 12! indexes (half reachable): counts equal
    SlidingPuzzleTests/TestFrontierSearch.cpp:71 (3x4 by transposition);
    byte and atomic cross-window marks."""
    w, h = (4, 3) if shape == "sliding_4x3" else (3, 4)
    p = psa.sliding(w, h)
    with psa.Bfs(p, segment_bits=24, direction=psa.DIR_AUTO, cross_marks=cross) as b:
        counts, hashes = b.run(0)
    assert counts == golden["sliding_4x3"]["counts"]
    if shape in FIX:
        assert [str(x) for x in hashes] == FIX[shape]["hashes"]
    if cross == 1:
        # top-down, finalize and expansion of each tile fused (9!-state tiles
        # inside one 12!-index segment)
        with psa.Bfs(p, cross_marks=cross, fused_layers=1) as b:
            assert b.run(0) == (counts, hashes)


def test_fused_layers_refused_outside_its_scope(ps):
    """fused_layers needs PS_CROSS_ATOMIC top-down batched searches whose
    tiles lie inside one segment."""
    p = psa.pancake(9)
    for bad in (dict(cross_marks=psa.CROSS_BYTES), dict(direction=psa.DIR_AUTO), dict(skip_visited=1),
                dict(mode=psa.BFS_PLAIN), dict(tile_bits_log2=-1), dict(segment_bits=12, tile_bits_log2=13)):
        with pytest.raises(psa.PSError):
            psa.Bfs(p, **{"cross_marks": psa.CROSS_ATOMIC, "fused_layers": 1, **bad})


def test_hanoi16_full(ps):
    """4-peg Hanoi, 16 disks, 4^16 states: every state is reached once; the
    search with byte cross-window marks and the batched search with atomic
    marks (2^20-state segments, the tuned shape the benchmark times) agree in
    every layer count and frontier hash, and equal the oracle's 162 layers
    (tests/golden/make_oracle_fixtures.py hanoi_16)."""
    with psa.Bfs(psa.hanoi(16), segment_bits=24) as b:
        counts, hashes = b.run(0)
        st = b.stats()
    assert sum(counts) == 4 ** 16
    assert st.expanded == 4 ** 16
    assert counts[:11] == [1, 3, 6, 12, 30, 30, 66, 96, 126, 210, 330]  # shared prefix with the 21-disk vector
    with psa.Bfs(psa.hanoi(16), segment_bits=20, block_threads=256, cross_marks=psa.CROSS_ATOMIC, graph=1) as b:
        assert b.run(0) == (counts, hashes)
    with psa.Bfs(psa.hanoi(16), segment_bits=20, cross_marks=psa.CROSS_ATOMIC, fused_layers=1) as b:
        assert b.run(0) == (counts, hashes)
    if "hanoi_16" in FIX:
        assert counts == FIX["hanoi_16"]["counts"]
        assert [str(h) for h in hashes] == FIX["hanoi_16"]["hashes"]


@pytest.mark.parametrize("vid,disks", [("hanoi_17_120", 17), ("hanoi_18_120", 18), ("hanoi_19_120", 19)])
def test_hanoi_deep_prefix(ps, golden, vid, disks):
    """4-peg Hanoi with 17, 18 and 19 disks, the whole 4^n index space in HBM
    (three bit arrays of 2 GiB, 8 GiB and 32 GiB): the first 121 layers equal
    the reference's vertical Hanoi search (HanoiTowersTests/HanoiSearchTests.cpp
    Search17..Search19, maxSteps 120).  The 17-disk vector leaves the others
    where the largest disk first moves, so moves of every disk are exercised."""
    r = golden[vid]
    print(f"hanoi {disks} disks, {len(r['counts'])} layers", flush=True)
    with psa.Bfs(psa.hanoi(disks), segment_bits=24, cross_marks=psa.CROSS_ATOMIC,
                 max_layers=len(r["counts"])) as b:
        counts, _ = b.run(0)
        st = b.stats()
    assert counts == r["counts"], r["source"]
    assert st.expanded == sum(r["counts"][:-1])  # every layer but the last is expanded
