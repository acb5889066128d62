"""Builds the native parts of puzzlesearch_amd in-tree.

  libpuzzlesearch_amd.so   HIP kernels + the C ABI (include/puzzlesearch_amd.h),
                           compiled with hipcc for gfx950
  libpuzzlesearch_host.so  the C++ mirror of the reference's Puzzle/PuzzleGpu
                           plugin interface and drivers, built on the C ABI

Usage: python -m puzzlesearch_amd.build [--jobs N] [--only amd|host]
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
OUT = os.path.join(PKG, "lib")
OBJ = os.path.join(PKG, "_build")
INCLUDE = os.path.join(ROOT, "include")

ARCH = os.environ.get("PS_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

HIP_FLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-variable",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}", f"-I{CSRC}",
]

AMD_LIB = os.path.join(OUT, "libpuzzlesearch_amd.so")
HOST_LIB = os.path.join(OUT, "libpuzzlesearch_host.so")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the packed-fp32 feature name; that line is noise
    lines = [ln for ln in r.stdout.splitlines() if "packed-fp32-ops" not in ln]
    return r.returncode, "\n".join(lines)


def _compile_hip(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    rc, log = _run([HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj])
    return src, obj, rc, log


def build_amd(jobs):
    os.makedirs(OUT, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    objs = []
    failed = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, obj, rc, log in ex.map(_compile_hip, srcs):
            if log.strip():
                print(log, file=sys.stderr)
            if rc != 0:
                failed.append(src)
            objs.append(obj)
    if failed:
        raise RuntimeError("hipcc failed for: " + ", ".join(failed))
    tmp = AMD_LIB + ".tmp"
    rc, log = _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
    if rc != 0:
        print(log, file=sys.stderr)
        raise RuntimeError("link of libpuzzlesearch_amd.so failed")
    os.replace(tmp, AMD_LIB)
    return AMD_LIB


def build_host(jobs):
    """C++ mirror of the reference interface; plain g++ over the C ABI."""
    if not os.path.isdir(HOST):
        return None
    srcs = sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp"))
    if not srcs:
        return None
    os.makedirs(OBJ, exist_ok=True)
    flags = ["-O2", "-std=c++17", "-fPIC", "-Wall", f"-I{INCLUDE}", f"-I{HOST}"]

    def one(src):
        obj = os.path.join(OBJ, "host_" + os.path.basename(src) + ".o")
        rc, log = _run([CXX] + flags + ["-c", src, "-o", obj])
        return src, obj, rc, log

    objs, failed = [], []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, obj, rc, log in ex.map(one, srcs):
            if log.strip():
                print(log, file=sys.stderr)
            if rc != 0:
                failed.append(src)
            objs.append(obj)
    if failed:
        raise RuntimeError("g++ failed for: " + ", ".join(failed))
    tmp = HOST_LIB + ".tmp"
    rc, log = _run([CXX, "-shared", "-o", tmp] + objs + [f"-L{OUT}", "-lpuzzlesearch_amd",
                                                           "-Wl,-rpath,$ORIGIN", "-lpthread"])
    if rc != 0:
        print(log, file=sys.stderr)
        raise RuntimeError("link of libpuzzlesearch_host.so failed")
    os.replace(tmp, HOST_LIB)
    return HOST_LIB


HOST_TESTS_SRC = os.path.join(ROOT, "tests", "cpp", "host_tests.cpp")
HOST_TESTS_BIN = os.path.join(ROOT, "tests", "cpp", "host_tests")


def build_host_tests():
    """The C++ test program of the host mirror (tests/cpp/host_tests)."""
    if not os.path.exists(HOST_TESTS_SRC):
        return None
    tmp = HOST_TESTS_BIN + ".tmp"
    rc, log = _run([CXX, "-O1", "-std=c++17", "-Wall", f"-I{INCLUDE}", f"-I{HOST}", HOST_TESTS_SRC, "-o", tmp,
                    f"-L{OUT}", "-lpuzzlesearch_host", "-lpuzzlesearch_amd", "-Wl,-rpath,$ORIGIN/../../puzzlesearch_amd/lib", "-lpthread"])
    if rc != 0:
        print(log, file=sys.stderr)
        raise RuntimeError("build of tests/cpp/host_tests failed")
    os.replace(tmp, HOST_TESTS_BIN)
    return HOST_TESTS_BIN


REF_TESTS_SRC = os.path.join(ROOT, "tests", "cpp", "ref_binding_tests.cpp")
REF_TESTS_BIN = os.path.join(ROOT, "tests", "cpp", "ref_binding_tests")
ORACLE = os.path.join(ROOT, "oracle")


def build_ref_binding_tests():
    """The C++ test program of the reference-signature binding
    (tests/cpp/ref_binding_tests): it checks the launchers against the CPU
    oracle, so it links the oracle library (test infrastructure), built first."""
    if not os.path.exists(REF_TESTS_SRC):
        return None
    rc, log = _run(["make", "-s", "-C", ORACLE, "liboracle.so"])
    if rc != 0:
        print(log, file=sys.stderr)
        raise RuntimeError("build of oracle/liboracle.so failed")
    tmp = REF_TESTS_BIN + ".tmp"
    rc, log = _run([CXX, "-O1", "-std=c++17", "-Wall", f"-I{INCLUDE}", f"-I{HOST}", f"-I{ORACLE}", REF_TESTS_SRC,
                    "-o", tmp, f"-L{OUT}", f"-L{ORACLE}", "-lpuzzlesearch_host", "-lpuzzlesearch_amd", "-loracle",
                    "-Wl,-rpath,$ORIGIN/../../puzzlesearch_amd/lib", "-Wl,-rpath,$ORIGIN/../../oracle", "-lpthread"])
    if rc != 0:
        print(log, file=sys.stderr)
        raise RuntimeError("build of tests/cpp/ref_binding_tests failed")
    os.replace(tmp, REF_TESTS_BIN)
    return REF_TESTS_BIN


def build(jobs=None, only=None):
    jobs = jobs or min(8, os.cpu_count() or 4)
    out = []
    if only in (None, "amd"):
        out.append(build_amd(jobs))
    if only in (None, "host"):
        h = build_host(jobs)
        if h:
            out.append(h)
            for t in (build_host_tests(), build_ref_binding_tests()):
                if t:
                    out.append(t)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--only", choices=["amd", "host"], default=None)
    args = ap.parse_args()
    for p in build(args.jobs, args.only):
        print("built", os.path.relpath(p, ROOT))


if __name__ == "__main__":
    main()
