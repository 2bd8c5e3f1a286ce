"""The C++ drop-in headers (include/ftk2/core/*.hpp, include/ndarray/ndarray.hh):
they compile as the reference's callers use them, and a program written against
them reproduces the reference's end-to-end cases on the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "test_examples_gpu.cpp")
BIN = os.path.join(ROOT, "tests", "cpp", "test_examples_gpu")


def test_headers_compile():
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-I", os.path.join(ROOT, "include"), SRC],
                   check=True)


def test_feature_element_layout():
    code = r'''
#include <ftk2/core/feature.hpp>
#include <cstddef>
static_assert(sizeof(ftk2::Simplex) == 48, "Simplex");
static_assert(sizeof(ftk2::FeatureElement) == 328, "FeatureElement");
static_assert(offsetof(ftk2::FeatureElement, track_id) == 248, "track_id");
int main() { return 0; }
'''
    p = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-x", "c++", "-I", os.path.join(ROOT, "include"), "-"],
                       input=code.encode(), capture_output=True)
    assert p.returncode == 0, p.stderr.decode()


@pytest.mark.gpu
def test_cpp_examples_on_gpu():
    assert os.path.exists(BIN), "build with __graft_entry__.build()"
    print("[cpp] running", BIN, flush=True)
    p = subprocess.run([BIN], capture_output=True, text=True, timeout=110)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "66418" in p.stdout and "6078" in p.stdout and "critical_point_2d: 788" in p.stdout
    assert "streaming:" in p.stdout  # execute_stream == execute through ftk::stream<>
    assert "streaming_retired:" in p.stdout  # the same with retirement after every slab
    assert "attributes_3d:" in p.stdout
    assert "extruded_2d:" in p.stdout  # ExtrudedSimplicialMesh over a regular base
    assert "multicomponent:" in p.stdout  # vector_var_name (the default input mode) and classification
