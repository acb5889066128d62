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


REF_BIN = os.path.join(ROOT, "tests", "cpp", "ref_binding_tests")


def _run_ref(flag):
    if not os.path.exists(REF_BIN):
        pytest.fail("tests/cpp/ref_binding_tests is not built (python -m puzzlesearch_amd.build)")
    r = subprocess.run([REF_BIN, flag], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    print(r.stderr)
    return r


def test_ref_binding_cpu_only():
    """Buffer<T> / TArray<T> with the interface of Common/Buffer.h and
    Common/Array.h."""
    r = _run_ref("--cpu")
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_ref_binding_launchers_on_gpu():
    """The reference-signature shim (Puzzles/GPU.h: element-count buffers and
    copies, ERR) and every launcher of Puzzles/PancakeGpu.h,
    Puzzles/SlidingPuzzleGpu.h and Puzzles/HanoiTowersGPU.h, called with the
    reference's arguments: slots equal the oracle's bit for bit."""
    r = _run_ref("--gpu")
    assert r.returncode == 0, r.stderr
    assert " 0 failures" in r.stdout
