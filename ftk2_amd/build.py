"""Build recipe for the in-tree native libraries.

libftk2h.so   HIP kernels + C ABI (include/ftk2h.h), gfx950 only.
libftk2synth.so  host-only synthetic field generators (bench and test inputs).

Run ``python -m ftk2_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

LIB_HIP = os.path.join(HERE, "libftk2h.so")
LIB_SYNTH = os.path.join(HERE, "libftk2synth.so")

# -ffp-contract=off: no product is fused into a following add unless written as
# fma(); the packed-fp32 feature is disabled for every HIP build.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip(jobs: int = 0) -> str:
    """Compile every HIP translation unit (in parallel) and link libftk2h.so."""
    import concurrent.futures as cf
    import glob
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    compile_flags = [f for f in HIP_FLAGS if f != "-shared"]
    objs = []

    def one(src):
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        _run([HIPCC, *compile_flags, "-c", src, "-o", obj])
        return obj

    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(one, srcs))
    tmp = LIB_HIP + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB_HIP)
    return LIB_HIP


def build_synth() -> str:
    tmp = LIB_SYNTH + ".tmp"
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
          os.path.join(CSRC, "synth.cpp"), "-o", tmp])
    os.replace(tmp, LIB_SYNTH)
    return LIB_SYNTH


CPP_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "test_examples_gpu.cpp")
CPP_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "test_examples_gpu")


def build_cpp_tests() -> str:
    """The C++ program written against the drop-in headers (include/ftk2/...)."""
    tmp = CPP_TEST_BIN + ".tmp"
    _run(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"), CPP_TEST_SRC,
          "-L", HERE, "-lftk2h", "-Wl,-rpath,$ORIGIN/../../ftk2_amd", "-o", tmp])
    os.replace(tmp, CPP_TEST_BIN)
    return CPP_TEST_BIN


def build_all() -> None:
    build_synth()
    build_hip()
    build_cpp_tests()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "hip": build_hip, "synth": build_synth, "cpp": build_cpp_tests}[w]()
