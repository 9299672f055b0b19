"""Device tree build (sfmm_tree_desc.on_device) against the host build.

The host build is pinned to the reference partitioner by the list digests and
level lists in test_tree.py; the device build must produce the same tree array
for array (equal FNV-1a digests over every array), so the plan and every
evaluation built on it are bitwise identical.
"""
import numpy as np
import pytest

from simplefmm_amd import api, workloads as W

pytestmark = pytest.mark.gpu


def _both(pos, box, max_level, cap):
    host = api.OctreePartitioner(pos, box).refine(max_level, cap)
    dev = api.OctreePartitioner(pos, box).refine(max_level, cap, on_device=True)
    return host, dev


def _duplicates(n):
    rng = np.random.default_rng(4)
    base = rng.random((7, 3))
    return np.ascontiguousarray(base[rng.integers(0, 7, n)])


CASES = [
    ("uniform", lambda: W.uniform_cube(20_000)[0], 20, 16, {}),
    ("plummer", lambda: W.plummer(200_000)[0], 20, 64, {}),
    ("clusters", lambda: W.clusters(50_000)[0], 20, 10, {}),
    ("sphere", lambda: W.sphere_shell(30_000)[0], 20, 8, {}),
    ("duplicates_max_level", lambda: _duplicates(5_000), 20, 16, {}),   # leaves at level 20 over cap
    ("cap0_level6", lambda: W.uniform_cube(3_000)[0], 6, 0, {}),
    ("max_level0", lambda: W.uniform_cube(500)[0], 0, 4, {}),           # one leaf
    ("two_points", lambda: W.uniform_cube(2)[0], 20, 1, {}),          # one point has a zero-width box
    ("depth0", lambda: W.uniform_cube(60)[0], 20, 64, {}),
    ("axis0_phi30", lambda: W.plummer(40_000)[0], 20, 32, {"vertical_axis": 0, "phideg": 30.0}),
]


@pytest.mark.parametrize("name,make,max_level,cap,boxkw", CASES, ids=[c[0] for c in CASES])
def test_device_tree_equals_host_tree(name, make, max_level, cap, boxkw):
    pos = make()
    print(f"case {name} n={len(pos)} max_level={max_level} cap={cap}", flush=True)
    host, dev = _both(pos, api.Box(pos, **boxkw), max_level, cap)
    assert dev.digest() == host.digest()
    for l in range(host.level() + 1):
        a, b = host.level_lists(l), dev.level_lists(l)
        for k in a:
            assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), (l, k)


def test_device_tree_explicit_box_and_outside_point():
    pos = W.uniform_cube(4_000)[0]
    box = api.Box(corner=[-0.01, -0.01, -0.01], width=1.02)
    host, dev = _both(pos, box, 20, 16)
    assert dev.digest() == host.digest()
    bad = pos.copy()
    bad[123] = [2.0, 0.5, 0.5]
    with pytest.raises(RuntimeError, match="outside the box"):
        api.OctreePartitioner(bad, box).refine(20, 16, on_device=True)


def test_device_tree_plan_evaluates_bitwise_like_host():
    pos, q = W.plummer(30_000)
    host, dev = _both(pos, api.Box(pos), 20, 32)
    out = []
    for part in (host, dev):
        fmm = api.FMM3D(part, 8, 1)
        U = np.zeros((len(pos), 4))
        fmm.rinv_potgrad(q, U)
        out.append(U)
    assert np.array_equal(out[0], out[1])


def test_device_tree_full_size_plummer():
    pos = W.plummer(1 << 22)[0]
    print(f"plummer n={len(pos)} cap=64", flush=True)
    host, dev = _both(pos, api.Box(pos), 20, 64)
    assert dev.digest() == host.digest()
