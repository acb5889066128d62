"""Runs the C++ test program of the host mirror (tests/cpp/host_tests.cpp):
the reference-interface classes (PancakeOptimizedGPU, SlidingTilePuzzleGpu,
SlidingTilePuzzleOptimized<w,h>, FourPegHanoi*GPU) and drivers
(InMemoryTwoBitBFS, InMemoryThreeBitBFS and their device-resident variants)
against the reference's own test expectations."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "host_tests")


def _run(flag):
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/host_tests is not built (python -m puzzlesearch_amd.build)")
    r = subprocess.run([BIN, flag], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    print(r.stderr)
    return r


def test_host_mirror_cpu_only():
    r = _run("--cpu")
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_host_mirror_on_gpu():
    r = _run("--gpu")
    assert r.returncode == 0, r.stderr
    assert " 0 failures" in r.stdout
