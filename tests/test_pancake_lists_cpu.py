"""CPU check of the cross-tile list arithmetic of the 10..12-pancake search
(puzzlesearch_amd/csrc/pancake_lists.h, compiled for the host): children
against the oracle's rank/unrank, list slots and capacities, and the apply
pass's enumeration of the lists reaching each tile (tests/cpp/pancake_lists_check.cpp)."""
import os
import subprocess

import pytest

import oracle_lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "pancake_lists_check.cpp")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    oracle_lib.lib()  # builds oracle/liboracle.so if needed
    out = str(tmp_path_factory.mktemp("pkl") / "pancake_lists_check")
    odir = os.path.join(ROOT, "oracle")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "puzzlesearch_amd", "csrc"), SRC,
                    "-o", out, "-L" + odir, "-loracle", "-Wl,-rpath," + odir], check=True)
    return out


@pytest.mark.parametrize("n,tiles", [(10, list(range(10))), (11, [0, 37, 109]), (12, [0, 777, 1319])])
def test_list_arithmetic(checker, n, tiles):
    r = subprocess.run([checker, str(n)] + [str(t) for t in tiles], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"pancake {n}: 0 failures" in r.stdout
