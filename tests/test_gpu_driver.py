"""The C++ drop-in (simplefmm_amd/cpp/fmm.hpp) through its main.cpp counterpart:
input and output files, against the Python binding of the same library."""
import hashlib
import os
import subprocess

import numpy as np
import pytest

from simplefmm_amd import api, workloads as W

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "simplefmm_amd", "cpp",
                      "sfmm_driver")


def _fnv1a(data: bytes) -> int:
    h = 1469598103934665603
    for b in data:
        h = ((h ^ b) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def _read_output(path):
    with open(path, "rb") as f:
        assert f.read(4) == b"SFMR"
        ver = int(np.frombuffer(f.read(4), np.int32)[0])
        n = int(np.frombuffer(f.read(8), np.int64)[0])
        nout = int(np.frombuffer(f.read(4), np.int32)[0])
        h = int(np.frombuffer(f.read(8), np.uint64)[0])
        out = np.frombuffer(f.read(), np.float64).reshape(n, nout)
    assert ver == 1
    return h, out


@pytest.mark.parametrize("qdim,device_tree", [(1, False), (3, False), (1, True)])
def test_driver_files_match_binding(tmp_path, qdim, device_tree):
    assert os.path.exists(DRIVER), "build the driver: make -C simplefmm_amd/cpp"
    pos, q = W.clusters(6000, seed=71)
    if qdim == 3:
        q = np.random.default_rng(3).uniform(-1, 1, (6000, 3))
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    W.write_harness_input(str(fin), pos, q)
    print(f"driver qdim={qdim} device_tree={device_tree}", flush=True)
    cmd = [DRIVER, "0", "48", "200", "1", "--in", str(fin), "--out", str(fout)] + (["--device-tree"] if device_tree else [])
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    h, out = _read_output(fout)
    assert h == _fnv1a(fin.read_bytes())
    part = api.OctreePartitioner(pos, api.Box(pos)).refine(20, 48)
    fmm = api.FMM3D(part, 8, ncharge=3)  # the driver's FMM3D<double, 3, 8, Vector, true>
    if qdim == 1:
        U = np.zeros((6000, 4))
        fmm.rinv_potgrad(q, U)
        assert np.array_equal(out, U)
    else:
        U = np.zeros((6000, 3))
        F = np.zeros((6000, 3))
        fmm.rinv(q, U)
        fmm.rinv(q, F, gradient=True)
        assert np.array_equal(out[:, :3], U) and np.array_equal(out[:, 3:], F)
        assert "Relative l2 error of field" in r.stdout
