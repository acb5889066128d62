"""GPU parity of the cross-tile list mode of the three-bit search of 10..12
pancakes (PS_CROSS_LISTS): per-layer counts and frontier hashes equal the
oracle's in batched (graph and eager) and plain (per-segment) execution, with
several segment sizes and block shapes, from the identity and from other
start states, with top-down layers only and with the late layers expanded
bottom-up (PS_DIR_AUTO: split layers, k_pkl_pull); the list bookkeeping is
left clean between runs; unsupported configurations are refused."""
import json
import math
import os

import pytest

import oracle_lib as O
import puzzlesearch_amd as psa

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "oracle_fixtures.json")) as f:
    FIX = json.load(f)

LISTS = dict(cross_marks=psa.CROSS_LISTS)
MODES = {
    "batched_graph": dict(mode=psa.BFS_BATCHED, graph=1),
    "batched_eager": dict(mode=psa.BFS_BATCHED, graph=0),
    "batched_seg20_256": dict(mode=psa.BFS_BATCHED, segment_bits=20, block_threads=256),
    "batched_seg12": dict(mode=psa.BFS_BATCHED, segment_bits=12),
    "plain": dict(mode=psa.BFS_PLAIN),
    "plain_seg20": dict(mode=psa.BFS_PLAIN, segment_bits=20),
    "plain_seg16_256": dict(mode=psa.BFS_PLAIN, segment_bits=16, block_threads=256),
    "auto_graph": dict(mode=psa.BFS_BATCHED, graph=1, direction=psa.DIR_AUTO),
    "auto_eager_512": dict(mode=psa.BFS_BATCHED, graph=0, direction=psa.DIR_AUTO, block_threads=512),
    "auto_seg20": dict(mode=psa.BFS_BATCHED, segment_bits=20, direction=psa.DIR_AUTO),
}


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("name", ["pancake_10", "pancake_11"])
def test_lists_against_oracle_fixture(ps, name, mode):
    rec = FIX[name]
    n = rec["puzzle"]["size"]
    print(f"{name} {mode}", flush=True)
    with psa.Bfs(psa.pancake(n), index_count=rec["index_count"], **LISTS, **MODES[mode]) as b:
        counts, hashes = b.run(rec["start"])
        st = b.stats()
    assert counts == rec["counts"]
    assert [str(h) for h in hashes] == rec["hashes"]
    if MODES[mode].get("direction") == psa.DIR_AUTO:
        # the last layers were found bottom-up: their parents were not expanded
        assert st.pull_words > 0
        assert st.expanded < math.factorial(n)
        assert st.expanded >= sum(counts[:len(counts) // 2])
    else:
        assert st.expanded == math.factorial(n)
    assert st.children == st.expanded * (n - 1)
    assert st.list_entries == st.expanded * (n - 9)
    if MODES[mode]["mode"] == psa.BFS_PLAIN:
        assert st.apply_launches >= len(counts) - 1
    else:  # one fused gather/finalize/expand kernel per layer
        assert st.apply_launches == 0 and st.finalize_launches == 0
        assert st.expand_launches >= len(counts)


@pytest.mark.parametrize("direction", [psa.DIR_TOP_DOWN, psa.DIR_AUTO])
@pytest.mark.parametrize("start", [1, 1234567, math.factorial(10) - 1])
def test_lists_other_starts(ps, start, direction):
    """Starts other than the identity: the oracle's counts and hashes."""
    op = O.pancake(10)
    oc, oh = O.bfs3(op, start)
    with psa.Bfs(psa.pancake(10), direction=direction, **LISTS) as b:
        assert b.run(start) == (oc, oh)


@pytest.mark.parametrize("direction", [psa.DIR_TOP_DOWN, psa.DIR_AUTO])
def test_lists_graph_replays_and_reruns(ps, direction):
    """Graph replays, eager runs and runs cut short by max_layers (before and
    inside the bottom-up layers) leave the lists and the visited states clean:
    every later run equals the first."""
    p = psa.pancake(11)
    d = dict(direction=direction, **LISTS)
    with psa.Bfs(p, graph=1, **d) as g, psa.Bfs(p, graph=0, **d) as e, \
            psa.Bfs(p, max_layers=5, **d) as m, psa.Bfs(p, max_layers=11, **d) as m2:
        ref = e.run(0)
        for _ in range(3):
            assert g.run(0) == ref
            cut = m.run(0)
            assert cut == (ref[0][:5], ref[1][:5])
            cut = m2.run(0)
            assert cut == (ref[0][:11], ref[1][:11])
            assert e.run(0) == ref
            psa.sync()
    assert ref[0] == FIX["pancake_11"]["counts"]
    assert [str(h) for h in ref[1]] == FIX["pancake_11"]["hashes"]


def test_lists_step_api(ps):
    rec = FIX["pancake_10"]
    with psa.Bfs(psa.pancake(10), **LISTS) as b:
        b.reset(0)
        counts, hashes = [1], []
        while True:
            c, h = b.step()
            if c == 0:
                break
            counts.append(c)
            hashes.append(h)
    assert counts == rec["counts"]
    assert [str(h) for h in hashes] == rec["hashes"][1:]


@pytest.mark.parametrize("seg,mode,direction", [(0, psa.BFS_BATCHED, 0), (22, psa.BFS_BATCHED, 0),
                                                (20, psa.BFS_PLAIN, 0), (22, psa.BFS_BATCHED, 1),
                                                (0, psa.BFS_BATCHED, 1)])
def test_pancake12_lists_full(ps, seg, mode, direction):
    """12-pancake, 12! states: counts and hashes equal the oracle fixture;
    sum = 12!, radius = P(12) = 14."""
    rec = FIX["pancake_12"]
    print(f"pancake_12 lists seg={seg} mode={mode} direction={direction}", flush=True)
    with psa.Bfs(psa.pancake(12), index_count=math.factorial(12), segment_bits=seg, mode=mode,
                 direction=direction, **LISTS) as b:
        counts, hashes = b.run(0)
    assert sum(counts) == math.factorial(12)
    assert len(counts) == 15
    assert counts == rec["counts"]
    assert [str(h) for h in hashes] == rec["hashes"]


def test_pancake12_lists_reference_sizing(ps):
    """PancakeOptimized's 2^29 index space (tiles past 12! hold no states)."""
    rec = FIX["pancake_12"]
    with psa.Bfs(psa.pancake(12), segment_bits=24, **LISTS) as b:
        counts, hashes = b.run(0)
    assert counts == rec["counts"]
    assert [str(h) for h in hashes] == rec["hashes"]


@pytest.mark.parametrize("puzzle,opts", [
    (psa.pancake(9), {}),
    (psa.pancake(13), dict(segment_bits=29)),
    (psa.sliding(4, 3), {}),
    (psa.hanoi(10), {}),
    (psa.pancake(10), dict(direction=psa.DIR_AUTO, mode=psa.BFS_PLAIN)),
    (psa.pancake(10), dict(direction=psa.DIR_BOTTOM_UP)),
    (psa.pancake(10), dict(direction=psa.DIR_AUTO, block_threads=256)),
    (psa.pancake(10), dict(tile_bits_log2=16)),
    (psa.pancake(10), dict(index_count=1000)),
    (psa.pancake(10), dict(bits=2)),
])
def test_lists_refused(ps, puzzle, opts):
    with pytest.raises(psa.PSError) as e:
        psa.Bfs(puzzle, **LISTS, **opts)
    assert e.value.code == psa.PS_ERR_UNSUPPORTED


def test_lists_footprint(ps):
    n = 12
    base = psa.bfs_footprint(psa.pancake(n), index_count=math.factorial(n), mode=psa.BFS_PLAIN,
                             cross_marks=psa.CROSS_ATOMIC)
    batched = psa.bfs_footprint(psa.pancake(n), index_count=math.factorial(n), cross_marks=psa.CROSS_ATOMIC)
    lists = psa.bfs_footprint(psa.pancake(n), index_count=math.factorial(n), **LISTS)
    plain = psa.bfs_footprint(psa.pancake(n), index_count=math.factorial(n), mode=psa.BFS_PLAIN, **LISTS)
    tiles = math.factorial(n) // math.factorial(9)
    one = tiles * (n - 9) * math.factorial(9) * 4 + tiles * 585 * 4 + tiles  # lists, counts, flags
    assert plain - base == one
    # a batched search without lists holds the list of tiles its expand
    # visits: 4 bytes per tile of at least 64 words
    nwords = (math.factorial(n) + 63) // 64
    assert batched - base == (nwords + 63) // 64 * 4
    # batched: lists of two layer parities, per-tile window flags, frontier
    # counts and the list of tiles a launch visits
    assert lists - base == 2 * one + 2 * tiles + 12 * tiles + 4 * tiles
