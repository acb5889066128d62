"""CPU emulation of the multi-search kernels of the fused batch mode
(k_multi_expand, k_multi_finalize, k_multi_layer_end in
puzzlesearch_amd/csrc/bfs_kernels.h over the puzzle table of
csrc/multi_kinds.h), compiled unchanged for the host over the SIMT emulator of
tests/cpp/simt with AddressSanitizer (tests/cpp/multi_emu.cpp).

Each mix runs step by step the way csrc/batch.hip drives the fused mode: job
arrays allocated with the exact sizes ps_bfs_create gives them, the job table
rebuilt whenever a job retires, three launches per step.  The emulator aborts
on any access outside an allocation (ASan, LDS beyond the launch's request
poisoned), on a block barrier or wave operation not reached by every lane, on
a deadlock and on a launch that runs past its time limit; every job's layer
counts and frontier hashes must equal the oracle's InMemoryThreeBitBFS
(Common/InMemoryThreeBitBFS.cpp:4-48).  The emulator itself is checked by
running the kernels to those exact results."""
import os
import subprocess

import pytest

import oracle_lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "multi_emu.cpp")


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    oracle_lib.lib()  # builds oracle/liboracle.so if needed
    out = str(tmp_path_factory.mktemp("emu") / "multi_emu")
    odir = os.path.join(ROOT, "oracle")
    subprocess.run(["g++", "-O1", "-g", "-std=c++20", "-fsanitize=address", "-fno-omit-frame-pointer",
                    "-fno-extern-tls-init",
                    "-I" + os.path.join(ROOT, "tests", "cpp", "simt"),
                    "-I" + os.path.join(ROOT, "puzzlesearch_amd", "csrc"),
                    "-I" + os.path.join(ROOT, "include"), "-I" + odir, SRC, "-o", out,
                    "-L" + odir, "-loracle", "-Wl,-rpath," + odir, "-lpthread"], check=True)
    return out


@pytest.mark.parametrize("args", [
    # pancake, sliding and Hanoi side by side, one segment each
    ["pancake:8", "sliding:2:4", "hanoi:6", "hanoi:4"],
    # segments of 2^12 states: segment counts, touched bits and per-segment tile liveness
    ["--segment-bits", "12", "pancake:8", "hanoi:7", "sliding:2:4"],
    # more blocks than tiles in flight, jobs retiring at different steps
    ["--grid", "6", "pancake:8", "sliding:2:4", "hanoi:5", "hanoi:7"],
])
def test_multi_kernels_emulated(emu, args):
    r = subprocess.run([emu] + args, capture_output=True, text=True, timeout=600)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "SIMT:" not in r.stderr
    assert " 0 failures" in r.stdout
