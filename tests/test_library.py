"""The C-ABI library loads and exports every entry point include/sfmm.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

from simplefmm_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "sfmm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sfmm_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    assert "sfmm_plan_create" in names and "sfmm_evaluate_device" in names
    assert len(names) >= 14


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(api.LIB_PATH)
    for name in _declared():
        assert hasattr(L, name), name


def test_binding_covers_header():
    assert set(_declared()) <= set(api.SIGNATURES)


def test_last_error_is_a_string():
    L = api.lib()
    assert isinstance(L.sfmm_last_error(), bytes)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU driver is present")
def test_device_tree_build_fails_loudly_without_a_gpu():
    from simplefmm_amd import workloads as W
    pos = W.uniform_cube(1000)[0]
    with pytest.raises(api.SfmmError, match="sfmm error 2"):
        api.OctreePartitioner(pos, api.Box(pos)).refine(20, 16, on_device=True)


def test_cpp_dropin_compiles():
    # the drop-in header (simplefmm_amd/cpp/fmm.hpp) with the calls code written for
    # the reference makes: self_generator constructor, comm(), slist(), rinv / N2N
    # with a lambda predicate (tests/dropin_api.cpp; run on the GPU by test_gpu_api.py)
    import subprocess
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-I",
                    os.path.join(ROOT, "simplefmm_amd", "cpp"), os.path.join(ROOT, "tests", "dropin_api.cpp")],
                   check=True, timeout=120)
