"""Shared helpers for the tests: golden fixtures and seeded inputs."""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np

from simplefmm_amd import workloads as W

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden_cases(mode: str | None = None, self_lists: bool | None = False, max_n: int | None = None):
    """Fixture names, optionally by mode, by whether the reference ran with a
    self generator (None: either), and by size."""
    out = []
    for f in sorted(os.listdir(GOLDEN)):
        if f.endswith(".json"):
            meta = json.load(open(os.path.join(GOLDEN, f)))
            if mode is not None and meta["mode"] != mode:
                continue
            if self_lists is not None and bool(meta.get("self_lists")) != self_lists:
                continue
            if max_n is not None and meta["n"] > max_n:
                continue
            out.append(meta["case"])
    return out


def golden_self_lists(meta, pos):
    """The self generator lists a self-list fixture was made with."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLDEN, "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.self_lists(pos, meta["seed"])


def load_golden(case: str):
    meta = json.load(open(os.path.join(GOLDEN, case + ".json")))
    data = dict(np.load(os.path.join(GOLDEN, case + ".npz"), allow_pickle=False))
    gen, n, seed = meta["generator"], meta["n"], meta["seed"]
    if gen == "uniform_cube_vector":
        pos, q = W.uniform_cube(n, seed=seed, charge="vector")
    else:
        pos, q = getattr(W, gen)(n, seed=seed)
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(pos).tobytes())
    h.update(np.ascontiguousarray(q).tobytes())
    if h.hexdigest() != meta["input_sha256"]:
        raise AssertionError(f"{case}: regenerated inputs differ from the fixture's (generator drift)")
    return meta, data, pos, q


def rel_l2(a, b) -> float:
    """sqrt(sum |a-b|^2 / sum |b|^2), b the truth (main.cpp:99-111)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


def list_sha(arr) -> str:
    return hashlib.sha256(np.ascontiguousarray(arr, dtype=np.int64).tobytes()).hexdigest()
