"""Builds the gfx950 device backend and its C++ interface in place.

    python -m spyker_amd.build            # compile every source, link
    python -m spyker_amd.build --jobs 8

Each device source (csrc/) is compiled to an object under spyker_amd/_build/
by hipcc for gfx950 only, then linked into spyker_amd/libspyker_hip.so (the
C ABI).  The C++ interface (cpp/spyker.cpp, include/spyker_amd/spyker.h) is
host code over that ABI, linked into spyker_amd/libspyker_cpp.so, and its
test driver into tests/cpp/api_test.
"""

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libspyker_hip.so")
OUT_CPP = os.path.join(HERE, "libspyker_cpp.so")
API_TEST = os.path.join(ROOT, "tests", "cpp", "api_test")
OBJ = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # one fused multiply-add only where the source writes one
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
    "-I" + CSRC,
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def compile_one(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s%s" % (src, r.stdout, r.stderr))
    return obj


def build(jobs=None, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, len(srcs))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC"] + objs + ["-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s%s" % (r.stdout, r.stderr))
    os.replace(tmp, OUT)
    build_cpp()
    if verbose:
        print("built", OUT)
        print("built", OUT_CPP)
        print("built", API_TEST)
    return OUT


HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-I" + os.path.join(ROOT, "include")]


def _run(cmd, what):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("%s failed:\n%s%s" % (what, r.stdout, r.stderr))


def build_cpp():
    """The C++ interface library and its test driver (host code, linked against the ABI library)."""
    tmp = OUT_CPP + ".tmp"
    _run([HIPCC] + HOST_FLAGS + ["-shared", os.path.join(HERE, "cpp", "spyker.cpp"), "-L" + HERE, "-lspyker_hip",
                                 "-Wl,-rpath,$ORIGIN", "-o", tmp], "C++ interface")
    os.replace(tmp, OUT_CPP)
    _run([HIPCC] + HOST_FLAGS + [os.path.join(ROOT, "tests", "cpp", "api_test.cpp"), "-L" + HERE, "-lspyker_cpp",
                                 "-lspyker_hip", "-Wl,-rpath,$ORIGIN/../../spyker_amd", "-o", API_TEST],
         "C++ interface test driver")


def main(argv=None):
    p = argparse.ArgumentParser()
    p.add_argument("--jobs", type=int, default=None)
    a = p.parse_args(argv)
    build(a.jobs, verbose=True)


if __name__ == "__main__":
    sys.exit(main())
