"""Host code of the plan under AddressSanitizer + UndefinedBehaviorSanitizer.

The octree, Morton order, interaction/near lists (tree.cpp) and the operator
tables (tables.cpp) are compiled with g++ -fsanitize=address,undefined
together with tests/sanitize/host_checks.cpp, which builds trees over uniform,
clustered, planar, collinear and coincident point sets (one and four host
threads, with the reference's box rule and with an explicit box) and the
tables for p = 1..16; any sanitizer report or failed invariant fails the test.
CPU only: no device code is built.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "simplefmm_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_tree_and_tables_clean_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "host_checks")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", "-pthread", os.path.join(HERE, "sanitize", "host_checks.cpp"),
           os.path.join(CSRC, "tree.cpp"), os.path.join(CSRC, "tables.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=300)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert "ok (0 failed checks)" in r.stdout
