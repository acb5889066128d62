"""CPU check of the packed sliding-tile move generator (psd::Packed<W,H> of
puzzlesearch_amd/csrc/device_puzzles.h compiled for the host,
tests/cpp/packed_check.cpp): the vertical moves update only the W packed
values they rotate, from the relative order of those W tiles and the parity
bit of list position 1; every slot of random states of all eleven shapes, with
random used-operator bits, must equal the oracle's restatement of the
reference's unpack / rotate / pack kernel (Puzzles/SlidingPuzzleGpu.cu:108-270).
A dropped parity bit fails every shape (checked when the arithmetic was
written).  The same program checks the simple Lehmer-rank encoding
(psd::Sliding<N>), whose vertical moves change only the digits of the swapped
window, on nine shapes against the oracle's slots of the reference's rank /
swap / re-rank kernel; an off-by-one in the moved tile's new digit fails all
nine."""
import os
import subprocess

import pytest

import oracle_lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "packed_check.cpp")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    oracle_lib.lib()
    out = str(tmp_path_factory.mktemp("packed") / "packed_check")
    odir = os.path.join(ROOT, "oracle")
    subprocess.run(["g++", "-O2", "-std=c++20", "-fno-extern-tls-init",
                    "-I" + os.path.join(ROOT, "tests", "cpp", "simt"),
                    "-I" + os.path.join(ROOT, "puzzlesearch_amd", "csrc"),
                    "-I" + os.path.join(ROOT, "include"), "-I" + odir, SRC, "-o", out,
                    "-L" + odir, "-loracle", "-Wl,-rpath," + odir, "-lpthread"], check=True)
    return out


@pytest.mark.parametrize("seed", [1, 2])
def test_packed_slots_equal_oracle(checker, seed):
    r = subprocess.run([checker, "200000", str(seed)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 shapes failed" in r.stdout
