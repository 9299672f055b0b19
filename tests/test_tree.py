"""Host octree (libsfmm.so, no GPU needed) against the reference and the oracle.

The reference's per-level lists (partition.hpp:155-166) are pinned through the
SHA-256 digests the golden fixtures record from the reference itself; further
configurations are compared list by list with the oracle's recursive
restatement, and the near/far split is checked by brute force: every ordered
pair of leaves is covered exactly once, by the near list or by one M2L pair of
ancestors.
"""
import numpy as np
import pytest

from oracle import oracle as O
from simplefmm_amd import api, workloads as W
from tests.helpers import golden_cases, list_sha, load_golden

LIST_KEYS = ["c_node", "leaf2node", "idx_row", "idx_val", "ilist_row", "ilist_val", "clist_row", "clist_val",
             "plist", "nlist_row", "nlist_val"]
# the reference dump's names for the same lists
DUMP_KEYS = {"c_node": "c_node", "leaf2node": "leaf2node", "idx_row": "idx_row", "idx_val": "idx_val",
             "ilist_row": "ilist_row", "ilist_val": "ilist_val", "clist_row": "clist_row",
             "clist_val": "clist_val", "plist": "plist", "nlist_row": "nlist_row", "nlist_val": "nlist_val"}


def _tree(pos, cap, max_level=20, **box):
    return api.OctreePartitioner(pos, api.Box(pos, **box)).refine(max_level, cap)


@pytest.mark.parametrize("case", golden_cases())
def test_lists_match_reference_digests(case):
    meta, _, pos, _ = load_golden(case)
    part = _tree(pos, meta["cap"], meta["max_level"])
    assert part.level() == meta["depth"]
    for l in range(meta["depth"] + 1):
        ref = meta["levels"][str(l)]
        mine = part.level_lists(l)
        assert mine["n_node"] == ref["n_node"] and mine["n_leaf"] == ref["n_leaf"]
        for k in LIST_KEYS:
            assert list_sha(mine[k]) == ref[DUMP_KEYS[k]], f"level {l} list {k}"


CONFIGS = [
    ("uniform", dict(n=3000, cap=8, max_level=20)),
    ("uniform", dict(n=2, cap=4, max_level=20)),
    ("uniform", dict(n=50, cap=64, max_level=20)),      # depth 0: everything near
    ("uniform", dict(n=300, cap=64, max_level=20)),     # depth 1: still no far field
    ("uniform", dict(n=5000, cap=1, max_level=3)),      # fat leaves at max_level
    ("clusters", dict(n=4000, cap=5, max_level=20)),
    ("plummer", dict(n=4000, cap=16, max_level=20)),
    ("sphere", dict(n=3000, cap=10, max_level=20)),
]


def _pos(kind, n, seed):
    if kind == "uniform":
        return W.uniform_cube(n, seed=seed)[0]
    if kind == "clusters":
        return W.clusters(n, seed=seed)[0]
    if kind == "plummer":
        return W.plummer(n, seed=seed)[0]
    return W.sphere_shell(n, seed=seed)[0]


def _compare_with_oracle(pos, cap, max_level, vertical_axis=2, phideg=0.0):
    part = _tree(pos, cap, max_level, vertical_axis=vertical_axis, phideg=phideg)
    ot = O.Tree(pos, max_level, cap, vertical_axis, phideg)
    assert part.level() == ot.depth
    assert part.info().width == ot.box()[0]
    for l in range(ot.depth + 1):
        a, b = part.level_lists(l), ot.level(l)
        for k in LIST_KEYS:
            assert np.array_equal(a[k], b[k]), f"level {l} list {k}"
    return part


@pytest.mark.parametrize("kind,cfg", CONFIGS)
def test_lists_match_oracle(kind, cfg):
    _compare_with_oracle(_pos(kind, cfg["n"], 31), cfg["cap"], cfg["max_level"])


@pytest.mark.parametrize("axis,phideg", [(0, 0.0), (1, 30.0), (2, 45.0)])
def test_rotated_box_matches_oracle(axis, phideg):
    _compare_with_oracle(W.clusters(3000, seed=32)[0], 12, 20, axis, phideg)


def test_duplicate_points():
    pos = W.uniform_cube(500, seed=33)[0]
    pos = np.concatenate([pos, pos[:200]])  # coincident pairs end in the same leaf
    _compare_with_oracle(pos, 8, 6)


def _ancestor(part_lists, l, node, la):
    while l > la:
        node = part_lists[l]["plist"][2 * node]
        l -= 1
    return node


def test_every_leaf_pair_covered_once():
    pos = W.clusters(2500, seed=34)[0]
    part = _tree(pos, 6)
    D = part.level()
    L = [part.level_lists(l) for l in range(D + 1)]
    leaves = [(l, f, int(L[l]["leaf2node"][f])) for l in range(D + 1) for f in range(L[l]["n_leaf"])]
    il = []
    for l in range(D + 1):
        s = set()
        if l >= 2:
            row, val = L[l]["ilist_row"], L[l]["ilist_val"].reshape(-1, 4)
            for a in range(L[l]["n_node"]):
                for k in range(row[a], row[a + 1]):
                    s.add((a, int(val[k, 0])))
        il.append(s)
    near = {}
    for l in range(D + 1):
        row, val = L[l]["nlist_row"], L[l]["nlist_val"].reshape(-1, 2)
        for f in range(L[l]["n_leaf"]):
            near[(l, f)] = {(int(x), int(y)) for x, y in val[row[f]:row[f + 1]]}
    for la, fa, na in leaves:
        for lb, fb, nb in leaves:
            cover = int((lb, fb) in near[(la, fa)])
            for k in range(2, min(la, lb) + 1):
                if (_ancestor(L, la, na, k), _ancestor(L, lb, nb, k)) in il[k]:
                    cover += 1
            assert cover == 1, f"leaf pair {(la, fa)} {(lb, fb)} covered {cover} times"


def test_tree_errors():
    pos = W.uniform_cube(100, seed=35)[0]
    with pytest.raises(api.SfmmError):
        api.OctreePartitioner(pos, api.Box(pos)).refine(21, 8)
    with pytest.raises(api.SfmmError):  # a point outside an explicit box (box.hpp:60 asserts)
        api.OctreePartitioner(pos, api.Box(corner=(0.2, 0.2, 0.2), width=0.5)).refine(20, 8)
    flat = pos.copy()
    flat[:, 2] = 0.5  # no vertical extent: the reference's box doubling never terminates
    with pytest.raises(api.SfmmError):
        api.OctreePartitioner(flat, api.Box(flat)).refine(20, 8)
    with pytest.raises(api.SfmmError):
        api.OctreePartitioner(pos[:1], api.Box(pos[:1])).refine(20, 8)
    bad = pos.copy()
    bad[3, 1] = np.nan
    with pytest.raises(api.SfmmError):
        api.OctreePartitioner(bad, api.Box(corner=(0, 0, 0), width=1.0)).refine(20, 8)


def test_near_sources_are_the_near_list():
    # sfmm_tree_near_sources: the particles N2N visits for target i (fmm.hpp:779-785)
    pos, _ = W.plummer(3000, seed=82)
    part = api.OctreePartitioner(pos, api.Box(pos)).refine(20, 16)
    t = O.Tree(pos, 20, 16)
    for i in range(0, 3000, 211):
        src = part.near_sources(i)
        assert i in src and len(set(src)) == len(src)
        assert sorted(src) == [j for j in range(3000) if t.is_neighbor(i, j)]
    with pytest.raises(api.SfmmError):
        part.near_sources(3000)
