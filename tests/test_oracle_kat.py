"""CPU oracle against the reference's unit-test known answers (kat_data.py)
and cross-checks between independent restatements:
  * the 4-peg Hanoi linked-list kernel form vs the bit-plane kernel form, on
    every state of small instances (two different restatements of
    Puzzles/HanoiTowersGPU.cu);
  * the packed sliding-tile kernel form (rotate/pack/unpack of
    Puzzles/SlidingPuzzleGpu.cu) vs moving the blank on the plain board and
    re-packing with the CPU class's Pack/Unpack
    (Puzzles/SlidingTilePuzzleOptimized.cpp);
  * PuzzleGpu's pack/unpack of (index, op) (arithmetic shift keeps ~0).
"""
import itertools

import numpy as np
import pytest

import oracle_lib as O
from kat_data import (COMPACT_KATS, EXPAND_KATS, HANOI_INDEX_KATS, PANCAKE15_SEGMENTS, RANK_KATS,
                      SLIDING_CPU_KATS)


def _ints(s):
    return [int(x) for x in s.split()]


@pytest.mark.parametrize("perm,expect", COMPACT_KATS)
def test_compact(oracle, perm, expect):
    a = np.array(_ints(perm), dtype=np.int32)
    O.lib().or_perm_compact(a, len(a))
    assert a.tolist() == _ints(expect)
    O.lib().or_perm_uncompact(a, len(a))
    assert a.tolist() == _ints(perm)


@pytest.mark.parametrize("perm,expect", RANK_KATS)
def test_rank(oracle, perm, expect):
    a = np.array(_ints(perm), dtype=np.int32)
    assert O.lib().or_perm_rank(a, len(a)) == expect


@pytest.mark.parametrize("kat", EXPAND_KATS, ids=[k[0] for k in EXPAND_KATS])
def test_expand_kat_gpu_layout(oracle, kat):
    src, spec, state, expected = kat
    p = O.from_spec(spec)
    idx = O.parse(p, state)
    assert O.to_string(p, idx) == state
    ops = O.ops_count(p)
    mask = (1 << ops) - 1
    slots = O.expand_slots(p, [idx << ops])
    for k, (child, op) in enumerate(expected):
        v = int(slots[k])
        if child is None:
            assert v == O.INVALID, (src, k)
            assert v & mask == op
            continue
        assert v != O.INVALID, (src, k)
        assert O.to_string(p, v >> ops) == child, (src, k)
        if op is None:
            assert v & mask != mask
        else:
            assert v & mask == op, (src, k)


def test_pancake15_segments(oracle):
    p = O.pancake(15)
    ch, _ = O.expand_one(p, 0)
    assert [c >> 29 for c in ch] == PANCAKE15_SEGMENTS


@pytest.mark.parametrize("kat", SLIDING_CPU_KATS)
def test_sliding_cpu_order(oracle, kat):
    src, w, h, state, expected = kat
    p = O.sliding(w, h)
    ch, _ = O.expand_one(p, O.parse(p, state))
    assert [O.to_string(p, c) for c in ch[:2]] == expected, src


@pytest.mark.parametrize("kat", HANOI_INDEX_KATS)
def test_hanoi_index(oracle, kat):
    src, disks, state, idx = kat
    p = O.hanoi(disks)
    assert O.parse(p, state) == idx, src
    assert O.to_string(p, idx) == state


def test_hanoi_cpu_opt_symmetry_kat(oracle):
    # PuzzlesTests/FourPegHanoiTest.cpp:178-201 (FourPegHanoiOptimized, CPU order)
    p = O.hanoi(7, True)
    idx = O.parse(p, "3 4 6 0 1 5 7 0 2 0")
    ch = np.zeros(16, dtype=np.uint64)
    ops = np.zeros(16, dtype=np.int32)
    n = O.lib().or_hanoi_expand_cpu_opt(7, 1, idx, 0, ch, ops)
    got = [(O.to_string(p, int(c)), int(o)) for c, o in zip(ch[:n], ops[:n])]
    assert got == [("1 3 4 6 0 5 7 0 2 0", 0), ("2 3 4 6 0 1 5 7 0 0", 0), ("4 6 0 1 5 7 0 3 0 2", 2),
                   ("3 4 6 0 5 7 0 1 2 0", 2), ("3 4 6 0 5 7 0 2 0 1", 3), ("3 4 6 0 1 5 7 0 2 0", 2)]


@pytest.mark.parametrize("disks,sym", [(1, False), (3, False), (3, True), (5, True), (6, False), (7, True)])
def test_hanoi_two_restatements_agree(oracle, disks, sym):
    n = 4 ** disks
    idx = np.arange(n, dtype=np.uint64)
    for op_bits in (0, 5, 9):
        packed = (idx << np.uint64(4)) | np.uint64(op_bits)
        a = O.expand_slots(O.hanoi(disks, sym, opt=False), packed)
        b = O.expand_slots(O.hanoi(disks, sym, opt=True), packed)
        assert (a == b).all()


def _board_move(cells, w, h, blank, target):
    c = list(cells)
    c[blank], c[target] = c[target], c[blank]
    return c


@pytest.mark.parametrize("w,h", [(2, 2), (3, 2), (4, 2), (3, 3), (5, 2)])
def test_packed_kernel_form_vs_board(oracle, w, h):
    """Kernel-form packed expansion == move the blank on the board and re-pack."""
    size = w * h
    p = O.sliding_packed(w, h)
    L = O.lib()
    rng = np.random.default_rng(11)
    perms = list(itertools.permutations(range(size))) if size <= 8 else [tuple(rng.permutation(size)) for _ in range(3000)]
    checked = 0
    for perm in perms:
        cells = np.array(perm, dtype=np.int32)
        r = L.or_packed_from_standard(w, h, cells)
        if r < 0:
            continue  # other parity class
        slots = O.expand_slots(p, [int(r) << 4])
        blank = perm.index(0)
        targets = [blank - w, blank + w, blank - 1, blank + 1]
        ok = [blank >= w, blank < size - w, blank % w > 0, blank % w < w - 1]
        for k in range(4):
            if not ok[k]:
                assert int(slots[k]) == O.INVALID
                continue
            child = np.array(_board_move(perm, w, h, blank, targets[k]), dtype=np.int32)
            expect = L.or_packed_from_standard(w, h, child)
            assert expect >= 0
            assert int(slots[k]) >> 4 == expect, (perm, k)
        checked += 1
    assert checked > 0


def test_split_keeps_invalid(oracle):
    idx, op = O.split_index_ops(4, [O.INVALID, (5 << 4) | 3])
    assert int(idx[0]) == O.INVALID and int(op[0]) == 15
    assert int(idx[1]) == 5 and int(op[1]) == 3


@pytest.mark.parametrize("n", [13, 14, 15])
def test_pancake_segment_modes_cover_full(oracle, n):
    """In-segment (ops 0..10) + cross-segment (ops 11..n-2) slots == full slots."""
    rng = np.random.default_rng(n)
    p = O.pancake(n)
    total = O.indexes_count(p)
    idx = rng.integers(0, total, size=200, dtype=np.uint64)
    # keep only valid segmented indexes (low 29 bits < 12!)
    idx = idx[(idx & np.uint64((1 << 29) - 1)) < np.uint64(479001600)]
    packed = idx << np.uint64(n - 1)
    full = O.expand_slots(p, packed).reshape(len(idx), n - 1)
    ins = O.expand_slots(p, packed, O.EXPAND_IN_SEGMENT).reshape(len(idx), 11)
    xs = O.expand_slots(p, packed, O.EXPAND_CROSS_SEGMENT).reshape(len(idx), n - 12)
    assert (np.concatenate([ins, xs], axis=1) == full).all()
    # in-segment children stay in the parent's 29-bit segment
    seg = idx >> np.uint64(29)
    assert ((ins >> np.uint64(n - 1 + 29)) == seg[:, None]).all()
