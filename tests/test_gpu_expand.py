"""GPU parity of ps_expand (PuzzleGpu::ExpandGpu) against the oracle: the
device output must equal the reference kernels' u64 slot layout bit for bit,
for every puzzle kind, random parents with random used-operator bits, empty
and ragged batch sizes, a full 2^21-parent batch (Common/Puzzle.h:18), and the
reference unit-test known answers."""
import math

import numpy as np
import pytest

import oracle_lib as O
import puzzlesearch_amd as psa
from kat_data import EXPAND_KATS, PANCAKE15_SEGMENTS

pytestmark = pytest.mark.gpu


def _oracle_puzzle(p):
    return O.puzzle(p.kind, p.size, p.height, p.flags)


def random_indexes(p, n, rng):
    """Valid indexes of puzzle p (the index encodings' value ranges)."""
    if p.kind == psa.PANCAKE:
        if p.size <= 12:
            return rng.integers(0, math.factorial(p.size), size=n, dtype=np.uint64)
        segs = math.prod(range(13, p.size + 1))
        seg = rng.integers(0, segs, size=n, dtype=np.uint64)
        low = rng.integers(0, math.factorial(12), size=n, dtype=np.uint64)
        return (seg << np.uint64(29)) | low
    if p.kind == psa.SLIDING:
        return rng.integers(0, math.factorial(p.size * p.height), size=n, dtype=np.uint64)
    if p.kind == psa.SLIDING_PACKED:
        L = O.lib()
        out = []
        size = p.size * p.height
        while len(out) < n:
            cells = rng.permutation(size).astype(np.int32)
            r = L.or_packed_from_standard(p.size, p.height, cells)
            if r >= 0:
                out.append(r)
        return np.array(out, dtype=np.uint64)
    return rng.integers(0, 4 ** p.size, size=n, dtype=np.uint64)


PUZZLES = ([psa.pancake(n) for n in range(2, 17)]
           + [psa.pancake(n, True) for n in (13, 15)]
           + [psa.sliding(w, h) for (w, h) in ((2, 1), (2, 2), (3, 2), (2, 3), (4, 2), (3, 3), (5, 2), (4, 3),
                                                (3, 4), (6, 2), (7, 2), (5, 3), (4, 4), (8, 2))]
           + [psa.sliding_packed(w, h) for (w, h) in ((2, 2), (3, 2), (4, 2), (5, 2), (6, 2), (7, 2), (8, 2),
                                                       (3, 3), (4, 3), (5, 3), (4, 4))]
           + [psa.hanoi(d, s, o) for d in (1, 2, 5, 10, 16, 17, 24) for s in (False, True) for o in (False, True)])


@pytest.mark.parametrize("p", PUZZLES, ids=repr)
def test_expand_random_parents(ps, p):
    rng = np.random.default_rng(p.kind * 1000 + p.size * 10 + p.height)
    ops = psa.puzzle_info(p).operators_count
    n = 3000
    idx = random_indexes(p, n, rng)
    op_bits = rng.integers(0, 1 << ops, size=n, dtype=np.uint64)
    op_bits[: n // 2] = 0  # half the parents with no used operators
    packed = (idx << np.uint64(ops)) | op_bits
    got = psa.expand(p, packed)
    want = O.expand_slots(_oracle_puzzle(p), packed)
    assert got.shape == want.shape
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, f"first mismatch at slot {bad[0]}: got {got[bad[0]]:#x} want {want[bad[0]]:#x}"


@pytest.mark.parametrize("n", [13, 14, 15, 16])
@pytest.mark.parametrize("mode", [psa.EXPAND_IN_SEGMENT, psa.EXPAND_CROSS_SEGMENT])
@pytest.mark.parametrize("inv", [False, True])
def test_expand_segment_modes(ps, n, mode, inv):
    p = psa.pancake(n, inv)
    rng = np.random.default_rng(n * 7 + mode)
    idx = random_indexes(p, 2000, rng)
    packed = (idx << np.uint64(n - 1)) | rng.integers(0, 1 << (n - 1), size=2000, dtype=np.uint64)
    got = psa.expand(p, packed, mode)
    want = O.expand_slots(_oracle_puzzle(p), packed, mode)
    assert (got == want).all()


@pytest.mark.parametrize("count", [0, 1, 63, 255, 256, 257, 1000])
def test_expand_ragged_sizes(ps, count):
    p = psa.pancake(12)
    rng = np.random.default_rng(count)
    packed = random_indexes(p, count, rng) << np.uint64(11)
    got = psa.expand(p, packed)
    want = O.expand_slots(_oracle_puzzle(p), packed)
    assert got.shape == (count * 11,)
    assert (got == want).all()


def test_expand_full_batch_pancake12(ps):
    """A full ExpandBuffer batch: 2^21 parents (Common/Puzzle.h:18)."""
    p = psa.pancake(12)
    n = 1 << 21
    rng = np.random.default_rng(2024)
    packed = random_indexes(p, n, rng) << np.uint64(11)
    got = psa.expand(p, packed)
    want = O.expand_slots(_oracle_puzzle(p), packed)
    assert (got == want).all()
    valid = got != np.uint64(psa.INVALID_INDEX)
    assert valid.all()  # no used-operator bits: every flip exists


@pytest.mark.parametrize("kat", EXPAND_KATS, ids=[k[0] for k in EXPAND_KATS])
def test_expand_known_answers(ps, kat):
    src, spec, state, expected = kat
    op = O.from_spec(spec)
    p = psa.Puzzle(op.kind, op.size, op.height, op.flags)
    ops = psa.puzzle_info(p).operators_count
    mask = (1 << ops) - 1
    idx = O.parse(op, state)
    got = psa.expand(p, [idx << ops])
    for k, (child, label) in enumerate(expected):
        v = int(got[k])
        if child is None:
            assert v == psa.INVALID_INDEX, (src, k)
            continue
        assert O.to_string(op, v >> ops) == child, (src, k)
        if label is not None:
            assert v & mask == label, (src, k)


def test_pancake15_child_segments(ps):
    p = psa.pancake(15)
    got = psa.expand(p, [0])
    assert [int(v) >> 14 >> 29 for v in got] == PANCAKE15_SEGMENTS


@pytest.mark.parametrize("n", [13, 14, 15, 16])
def test_xseg_postprocess(ps, n):
    rng = np.random.default_rng(n)
    segs = math.prod(range(13, n + 1))
    for seg in (0, 1, segs - 1, int(rng.integers(0, segs))):
        idx = rng.integers(0, math.factorial(12), size=5000, dtype=np.uint64).astype(np.uint32)
        want = idx.copy()
        O.lib().or_pancake_xseg_postprocess(n, seg, want, len(want))
        buf = psa.DeviceBuffer.from_array(idx)
        psa.check(psa.lib().ps_pancake_xseg_postprocess(n, seg, buf.ptr, len(idx), None))
        got = buf.download(np.uint32, len(idx))
        assert (got == want).all()


def test_expand_rejects_bad_mode(ps):
    with pytest.raises(psa.PSError):
        psa.expand(psa.pancake(9), [0], psa.EXPAND_IN_SEGMENT)
