"""GPU parity of the HanoiTowers vertical solver (csrc/hanoi_vertical.hip):

  * ps_hanoi_expand_in_segment (GpuExpandInSegment[WithoutSmallest]<size>,
    HanoiTowers/HanoiExpandGPU.cu:98-275) slot for slot against the oracle's
    slot form on random reduced states of 14..24 disks in every segment they
    fall in;
  * the device search against the oracle's HanoiSearch restatement (restored
    and reduced counts and the reduced frontier hash of every layer) for 10,
    12 and 13 disks with segments of 2^8 .. 2^32 states, so that many segments
    are allocated as the search reaches them, full-symmetry segments take the
    four-bit shortcut and cross-segment children are applied;
  * the reference's own vectors: HanoiSearch<17..24> with maxSteps 120
    (HanoiTowersTests/HanoiSearchTests.cpp:15-105), and the complete 16-disk
    search against the unreduced puzzle's 162 layer counts (oracle fixture);
  * a reused search gives the same result; a start that is not a reduced state
    and a segment beyond the budget are refused."""
import json
import os

import numpy as np
import pytest

import oracle_lib as O
import puzzlesearch_amd as psa

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _random_reduced(disks, k, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(k):
        pegs = [int(x) for x in rng.integers(0, 4, size=disks)]
        bottom = {p: max([d for d in range(disks) if pegs[d] == p], default=-1) for p in (1, 2, 3)}
        order = sorted((1, 2, 3), key=lambda p: -bottom[p])
        m = {0: 0, order[0]: 1, order[1]: 2, order[2]: 3}
        out.append(sum(m[p] << (2 * d) for d, p in enumerate(pegs)))
    return out


@pytest.mark.parametrize("disks", [14, 16, 17, 20, 24])
def test_expand_in_segment_slots(ps, disks):
    states = _random_reduced(disks, 3000, 100 + disks)
    for seg in sorted(set(s >> 32 for s in states)):
        part = np.array([s & 0xFFFFFFFF for s in states if s >> 32 == seg], np.uint32)
        for ws in (False, True):
            dev = psa.hanoi_expand_in_segment(disks, seg, part, without_smallest=ws)
            ref = O.hv_expand_in_segment_slots(disks, seg, part, without_smallest=ws)
            assert np.array_equal(dev, ref), (disks, seg, ws)


def test_expand_rejects_segment_outside(ps):
    idx = np.zeros(4, np.uint32)
    with pytest.raises(ps.PSError):
        psa.hanoi_expand_in_segment(16, 1, idx)
    with pytest.raises(ps.PSError):
        psa.hanoi_expand_in_segment(17, 4, idx)


@pytest.mark.parametrize("disks,sb", [(10, 32), (10, 16), (10, 12), (10, 8), (12, 24), (12, 16), (13, 20), (13, 14)])
def test_search_against_oracle(ps, disks, sb):
    o_rest, o_red, o_hash = O.hv_search(disks, 0, sb)
    with psa.HanoiSearch(disks, sb) as h:
        rest, red, hashes = h.run(0)
        st = h.stats()
    print(f"{disks} disks, 2^{sb} segments: {len(rest)} layers, {st['segments_allocated']} segments", flush=True)
    assert rest == o_rest
    assert red == o_red
    assert hashes == o_hash
    assert st["expanded"] == sum(red)  # the last layer is expanded too: its children are all visited


def test_hanoi_10_golden(ps, golden):
    with psa.HanoiSearch(10, 8) as h:
        rest, red, _ = h.run(0)
    assert rest == golden["hanoi_10"]["counts"]
    assert red == golden["hanoi_10_sym"]["counts"]


@pytest.mark.parametrize("disks", [17, 18, 19, 20, 21, 22, 23, 24])
def test_reference_search_vectors(ps, golden, disks):
    r = golden[f"hanoi_{disks}_120"]
    with psa.HanoiSearch(disks) as h:
        rest, red, _ = h.run(0, max_layers=len(r["counts"]))
        st = h.stats()
    print(f"HanoiSearch<{disks}> 120 steps: {st['device_ms']:.1f} ms device, {st['segments_allocated']} segments, "
          f"{st['listed_children']} listed children", flush=True)
    assert rest == r["counts"], r["source"]


def test_hanoi_16_complete(ps):
    with open(os.path.join(HERE, "golden", "oracle_fixtures.json")) as f:
        fx = json.load(f)["hanoi_16"]
    with psa.HanoiSearch(16) as h:
        rest, red, _ = h.run(0)
        st = h.stats()
        again = h.run(0)
    print(f"Hanoi-16 reduced search: {st['device_ms']:.1f} ms device, {sum(red)} reduced states", flush=True)
    assert rest == fx["counts"]
    assert sum(rest) == 4 ** 16
    assert again[0] == rest and again[1] == red


def test_refusals(ps):
    with psa.HanoiSearch(12, 16) as h:
        with pytest.raises(ps.PSError):
            h.run(2 << 2)  # disk 1 on peg 2 with peg 1 empty: not reduced
        with pytest.raises(ps.PSError):
            h.run(4 ** 12)
        assert h.run(0)[0] == O.hv_search(12, 0, 16)[0]
    # two segments of 2^16 states need 3 * 8 KiB each
    with psa.HanoiSearch(12, 16, hbm_budget=3 * 8192 + 100) as h:
        with pytest.raises(ps.PSError):
            h.run(0)
    with pytest.raises(ps.PSError):
        psa.HanoiSearch(12, 7)
