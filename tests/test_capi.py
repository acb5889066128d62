"""C ABI checks that need no GPU: the library loads, exports every entry point
include/puzzlesearch_amd.h declares, answers host-only puzzle queries like the
reference classes (compared with the oracle), and reports errors as status
codes instead of crashing."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O
import puzzlesearch_amd as psa

PUZZLES = [psa.pancake(n) for n in (2, 5, 9, 11, 12, 13, 16)] + \
          [psa.sliding(w, h) for (w, h) in ((2, 2), (3, 2), (3, 3), (4, 3), (4, 4), (2, 5))] + \
          [psa.sliding_packed(w, h) for (w, h) in ((2, 2), (3, 2), (4, 3), (4, 4), (7, 2), (5, 3))] + \
          [psa.hanoi(d, s, o) for d in (1, 10, 16, 24) for s in (False, True) for o in (False, True)]


def _as_oracle(p):
    return O.puzzle(p.kind, p.size, p.height, p.flags)


def test_library_built():
    assert os.path.exists(psa.LIB_PATH), "run python -m puzzlesearch_amd.build"


def test_exports_every_declared_symbol():
    L = psa.lib()
    declared = psa.declared_symbols()
    assert len(declared) > 30
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", psa.LIB_PATH], capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert set(declared) <= exported


@pytest.mark.parametrize("p", PUZZLES, ids=repr)
def test_puzzle_query_matches_reference_classes(p):
    info = psa.puzzle_info(p)
    op = _as_oracle(p)
    assert info.operators_count == O.ops_count(op)
    assert info.branching_factor == O.lib().or_branching_factor(ctypes.byref(op))
    assert info.indexes_count == O.indexes_count(op)
    assert info.has_odd_cycles == O.lib().or_has_odd_cycles(ctypes.byref(op))
    for mode in (psa.EXPAND_FULL, psa.EXPAND_IN_SEGMENT, psa.EXPAND_CROSS_SEGMENT):
        assert psa.expand_slots(p, mode) == O.slots_per_parent(op, mode)


def test_pancake12_index_space_quirk():
    # PancakeOptimized::IndexesCount() returns 2^29 for n = 12 (m_Size < 12 test)
    assert psa.puzzle_info(psa.pancake(12)).indexes_count == 1 << 29
    assert psa.puzzle_info(psa.pancake(11)).indexes_count == 39916800


@pytest.mark.parametrize("bad", [psa.pancake(1), psa.pancake(17), psa.sliding(5, 4), psa.sliding_packed(3, 4),
                                 psa.hanoi(25), psa.Puzzle(9, 3, 0, 0)])
def test_invalid_puzzles_report_errors(bad):
    info = psa.PuzzleInfo()
    rc = psa.lib().ps_puzzle_query(ctypes.byref(bad), ctypes.byref(info))
    assert rc == psa.PS_ERR_INVALID_ARGUMENT
    assert psa.lib().ps_last_error()
    assert psa.expand_slots(bad) == -1


def test_footprint_is_host_only():
    p = psa.pancake(12)
    b3 = psa.bfs_footprint(p)
    assert b3 >= 3 * (1 << 29) // 8
    b2 = psa.bfs_footprint(p, bits=2)
    assert 2 * (1 << 29) // 8 <= b2 < b3
    seg = psa.bfs_footprint(p, segment_bits=20)
    assert seg > b3  # plus per-segment bookkeeping


def test_calls_before_init_fail_cleanly():
    if psa.device_count() > 0:
        pytest.skip("a GPU is present; this checks the no-GPU error path")
    ptr = ctypes.c_void_p()
    rc = psa.lib().ps_alloc(ctypes.byref(ptr), 64)
    assert rc in (psa.PS_ERR_NOT_INITIALIZED, psa.PS_ERR_HIP)
    assert psa.lib().ps_init(0) == psa.PS_ERR_HIP
    assert b"no HIP device" in psa.lib().ps_last_error()


def test_c_abi_has_no_cxx_types():
    import re
    hdr = open(os.path.join(os.path.dirname(psa.PKG_DIR), "include", "puzzlesearch_amd.h")).read()
    code = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    code = re.sub(r"//[^\n]*", "", code)
    for token in ("std::", "torch", "template", "class "):
        assert token not in code
