"""Pins the oracle's HanoiTowers vertical solver (oracle/or_hanoi_vertical.c):

  * the search reproduces the reference's own vectors: 17 disks for all 121
    layers the reference test runs (HanoiTowersTests/HanoiSearchTests.cpp:15-23),
    the first 60 layers of 24 disks (:97-105), and Hanoi 10, whose restored
    counts are the plain puzzle's (PuzzlesTests/PuzzleSearchTests.cpp:354) and
    whose reduced counts are the symmetric puzzle's (:372), with segments of
    2^8 .. 2^32 states, so the four-bit shortcut of full-symmetry segments, the
    InSegmentOnly filter and the cross-segment marks all take part;
  * for 12 disks the reduced layers equal the symmetric plugin puzzle's
    three-bit BFS and the restored counts the plain puzzle's (or_bfs3), with the
    same reduced frontier hashes;
  * PegsMult equals, by brute force, the number of distinct states that
    permuting pegs 1..3 of a reduced state gives (all states of 6 disks);
  * InSegmentOnly never holds for a state with a child outside its segment;
  * the GPU kernels' slot form equals the full expansion slot for slot (the
    parent's id where a child leaves the segment), and in a full-symmetry
    segment the moves without the smallest disk plus its four positions are
    exactly the full expansion's children."""
import itertools

import numpy as np
import pytest

import oracle_lib as O


def test_hanoi_10_any_segment_size(golden):
    plain, sym = golden["hanoi_10"]["counts"], golden["hanoi_10_sym"]["counts"]
    ref = None
    for sb in (32, 20, 16, 12, 8):
        restored, reduced, hashes = O.hv_search(10, 0, sb)
        assert restored == plain, sb
        assert reduced == sym, sb
        ref = ref or hashes
        assert hashes == ref, sb


def test_search17_all_layers(golden):
    r = golden["hanoi_17_120"]
    restored, _, _ = O.hv_search(17, 0, 32, len(r["counts"]), with_hashes=False)
    assert restored == r["counts"], r["source"]


def test_search24_prefix(golden):
    r = golden["hanoi_24_120"]
    restored, _, _ = O.hv_search(24, 0, 32, 60, with_hashes=False)
    assert restored == r["counts"][:60], r["source"]


def test_hanoi_12_equals_plugin_bfs():
    restored, reduced, hashes = O.hv_search(12, 0, 16)
    pc, _ = O.bfs3(O.hanoi(12), 0)
    sc, sh = O.bfs3(O.hanoi(12, symmetry=True), 0)
    assert restored == pc
    assert reduced == sc
    assert hashes == sh


def _pegs(disks, index):
    return [(index >> (2 * d)) & 3 for d in range(disks)]


def _index(pegs):
    return sum(p << (2 * d) for d, p in enumerate(pegs))


def _canonical(disks, index):
    pegs = _pegs(disks, index)
    bottom = [max([d for d in range(disks) if pegs[d] == p], default=-1) for p in range(4)]
    return bottom[1] >= bottom[2] >= bottom[3]


def test_pegs_mult_brute_force():
    disks = 6
    checked = 0
    for index in range(1, 4 ** disks):
        if not _canonical(disks, index):
            continue
        pegs = _pegs(disks, index)
        orbit = set()
        for perm in itertools.permutations((1, 2, 3)):
            m = {0: 0, 1: perm[0], 2: perm[1], 3: perm[2]}
            orbit.add(_index([m[p] for p in pegs]))
        assert O.hv_pegs_mult(index) == len(orbit), index
        checked += 1
    assert checked > 600
    # the start, every disk on peg 0, is counted 1 by the search itself
    assert O.hv_pegs_mult(0) == 2


def test_in_segment_only_has_no_cross_children():
    rng = np.random.default_rng(5)
    disks, sb = 9, 8
    seen = 0
    for index in rng.integers(0, 4 ** disks, size=20000):
        index = int(index)
        if not _canonical(disks, index) or not O.hv_in_segment_only(index & 0xFF, sb):
            continue
        seg = index >> sb
        assert all((c >> sb) == seg for c in O.hv_expand(disks, index)), index
        seen += 1
    assert seen > 1000


def _random_reduced(disks, k, seed):
    """Random reduced states: random pegs, then pegs 1..3 relabelled in the
    order of their bottom disks."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(k):
        pegs = [int(x) for x in rng.integers(0, 4, size=disks)]
        bottom = {p: max([d for d in range(disks) if pegs[d] == p], default=-1) for p in (1, 2, 3)}
        order = sorted((1, 2, 3), key=lambda p: -bottom[p])
        m = {0: 0, order[0]: 1, order[1]: 2, order[2]: 3}
        out.append(_index([m[p] for p in pegs]))
    return out


@pytest.mark.parametrize("disks", [14, 17, 20, 24])
def test_slot_form_matches_full_expansion(disks):
    states = _random_reduced(disks, 600, disks)
    assert all(_canonical(disks, x) for x in states)
    for seg in sorted(set(int(s) >> 32 for s in states)):
        part = np.array([int(s) & 0xFFFFFFFF for s in states if int(s) >> 32 == seg], np.uint32)
        six = O.hv_expand_in_segment_slots(disks, seg, part)
        three = O.hv_expand_in_segment_slots(disks, seg, part, without_smallest=True)
        for i, lo in enumerate(part):
            index = (seg << 32) | int(lo)
            full = O.hv_expand(disks, index)
            want = [c & 0xFFFFFFFF if c >> 32 == seg else int(lo) for c in full]
            assert list(six[6 * i:6 * i + 6]) == want
            p0 = index & 3
            pairs = [k for k, (a, b) in enumerate([(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)])
                     if p0 not in (a, b)]
            assert list(three[3 * i:3 * i + 3]) == [want[k] for k in pairs]


def test_full_symmetry_four_bit_shortcut():
    # segments whose disks above the segment hold pegs 1 and 2: PegsMult = 6
    disks, sb = 8, 8
    checked = 0
    for seg in range(4 ** (disks - sb // 2)):
        if O.hv_pegs_mult(seg) != 6:
            continue
        for lo in range(0, 1 << sb, 7):
            index = (seg << sb) | lo
            if not _canonical(disks, index):
                continue
            full = set(c for c in O.hv_expand(disks, index) if c >> sb == seg)
            b4 = index & ~3
            short = {b4 | q for q in range(4)} - {index}
            p0 = index & 3
            for a, b in [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]:
                if p0 in (a, b):
                    continue
                k = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)].index((a, b))
                c = O.hv_expand(disks, index)[k]
                if c >> sb == seg:
                    short.add(c)
            full.discard(index)
            short.discard(index)
            assert short == full, index
            checked += 1
    assert checked > 100
