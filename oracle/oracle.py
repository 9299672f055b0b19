"""ctypes binding of the CPU oracle (liboracle.so) — TEST INFRASTRUCTURE ONLY.

The oracle is a CPU restatement of the reference FMM (icyspirit/SimpleFMM) used
as the checker for the HIP product.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

POTGRAD, VECPOT, CURL, POT = 0, 1, 2, 3
FAR, NEAR, BOTH = 1, 2, 3
_NOUT = {POTGRAD: 4, VECPOT: 3, CURL: 3, POT: 1}
_QDIM = {POTGRAD: 1, VECPOT: 3, CURL: 3, POT: 1}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int64)


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_tree_build.restype = ctypes.c_void_p
        L.oracle_tree_build.argtypes = [ctypes.c_int64, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.oracle_tree_free.argtypes = [ctypes.c_void_p]
        L.oracle_tree_depth.argtypes = [ctypes.c_void_p]
        L.oracle_tree_box.argtypes = [ctypes.c_void_p, _dp]
        L.oracle_tree_counts.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip]
        L.oracle_tree_copy.argtypes = [ctypes.c_void_p, ctypes.c_int] + [_ip] * 11
        L.oracle_fmm.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_int]
        L.oracle_fmm_self.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int,
                                      ctypes.c_int, _ip, _ip]
        L.oracle_direct.argtypes = [ctypes.c_int64, _dp, ctypes.c_int, _dp, ctypes.c_int64, _ip, _dp]
        L.oracle_is_neighbor.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64]
        L.oracle_is_neighbor.restype = ctypes.c_int
        L.oracle_Z.restype = ctypes.c_double
        L.oracle_Z.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.oracle_A.restype = ctypes.c_double
        L.oracle_A.argtypes = [ctypes.c_int, ctypes.c_int]
        L.oracle_wigner_R.restype = ctypes.c_double
        L.oracle_wigner_R.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.oracle_harmonic.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                      _dp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


class Tree:
    """Box(positions, vertical_axis, phideg) + OctreePartitioner.refine(max_level, cap)."""

    def __init__(self, pos, max_level=20, cap=64, vertical_axis=2, phideg=0.0):
        self.pos = np.ascontiguousarray(pos, dtype=np.float64)
        self.n = self.pos.shape[0]
        self.h = lib().oracle_tree_build(self.n, _d(self.pos), max_level, cap, vertical_axis, phideg)
        if not self.h:
            raise ValueError("oracle_tree_build failed")
        self.depth = lib().oracle_tree_depth(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_tree_free(self.h)
            self.h = None

    def box(self):
        out = np.zeros(4)
        lib().oracle_tree_box(self.h, _d(out))
        return out[0], out[1:].copy()

    def level(self, l: int) -> dict:
        c = np.zeros(6, dtype=np.int64)
        lib().oracle_tree_counts(self.h, l, _i(c))
        n_node, n_leaf, nidx, nil, ncl, nnl = (int(x) for x in c)
        out = {
            "n_node": n_node, "n_leaf": n_leaf,
            "c_node": np.zeros(n_node, np.int64), "leaf2node": np.zeros(n_leaf, np.int64),
            "idx_row": np.zeros(n_leaf + 1, np.int64), "idx_val": np.zeros(nidx, np.int64),
            "ilist_row": np.zeros((n_node + 1) if l >= 2 else 1, np.int64), "ilist_val": np.zeros(4 * nil, np.int64),
            "clist_row": np.zeros((n_node + 1) if l < self.depth else 1, np.int64),
            "clist_val": np.zeros(2 * ncl, np.int64),
            "plist": np.zeros(2 * n_node if l >= 1 else 0, np.int64),
            "nlist_row": np.zeros(n_leaf + 1, np.int64), "nlist_val": np.zeros(2 * nnl, np.int64),
        }
        keys = ["c_node", "leaf2node", "idx_row", "idx_val", "ilist_row", "ilist_val", "clist_row", "clist_val",
                "plist", "nlist_row", "nlist_val"]
        lib().oracle_tree_copy(self.h, l, *[_i(out[k]) for k in keys])
        return out

    def is_neighbor(self, i: int, j: int) -> bool:
        """OctreePartitioner::is_neighbor(i, j) (partition.hpp:484-488)."""
        return bool(lib().oracle_is_neighbor(self.h, int(i), int(j)))

    def fmm(self, p: int, q, mode=POTGRAD, which=BOTH, dense=False, self_lists=None):
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        if q.size != self.n * _QDIM[mode]:
            raise ValueError("charge array has the wrong size for this mode")
        out = np.zeros(self.n * _NOUT[mode])
        if self_lists is None:
            rc = lib().oracle_fmm(self.h, p, mode, _d(q), _d(out), which, int(dense))
        else:
            row, val = (np.ascontiguousarray(a, dtype=np.int64) for a in self_lists)
            rc = lib().oracle_fmm_self(self.h, p, mode, _d(q), _d(out), which, int(dense), _i(row), _i(val))
        if rc != 0:
            raise RuntimeError(f"oracle_fmm failed ({rc})")
        return out.reshape(self.n, _NOUT[mode]) if _NOUT[mode] > 1 else out


def direct(pos, q, targets, mode=POTGRAD):
    """Direct sum over all j != target (FP64)."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
    targets = np.ascontiguousarray(targets, dtype=np.int64)
    out = np.zeros(len(targets) * _NOUT[mode])
    rc = lib().oracle_direct(pos.shape[0], _d(pos), mode, _d(q), len(targets), _i(targets), _d(out))
    if rc != 0:
        raise RuntimeError("oracle_direct failed")
    return out.reshape(len(targets), _NOUT[mode]) if _NOUT[mode] > 1 else out


def Z(n, m, theta):
    return lib().oracle_Z(n, m, theta)


def A(n, m):
    return lib().oracle_A(n, m)


def wigner_R(n, a, b, theta):
    return lib().oracle_wigner_R(n, a, b, theta)


def harmonic(p, x, y, z, grad=0):
    """Complex (p+1)^2 vector rho^n Z(n,m) e^{-im phi} of (x,y,z), or its d/dx_{grad-1}."""
    out = np.zeros(2 * (p + 1) ** 2)
    lib().oracle_harmonic(p, x, y, z, grad, _d(out))
    return out[0::2] + 1j * out[1::2]


def rel_l2(a, b):
    """Relative L2 as main.cpp:99-111: sqrt(sum |a-b|^2 / sum |b|^2) (b is the truth)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))
