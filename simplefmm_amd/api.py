"""ctypes binding of libsfmm.so and a Python mirror of the reference API.

The reference (icyspirit/SimpleFMM) is a header-only C++ library driven as in
its main.cpp:38-66:

    Box<double,3,Vector> box(positions, vertical_axis);
    OctreePartitioner<double,Vector> partitioner(positions, box);
    partitioner.refine(max_level, max_particles_per_node);
    FMM3D<double,3,p,Vector,true> fmm(partitioner, MPI_COMM_WORLD);
    fmm.rinv(Q, U);  fmm.rinv<true>(Q, U);

`Box`, `OctreePartitioner` and `FMM3D` below keep those names, argument
meanings and the accumulate-into-U semantics; the work runs in libsfmm.so
(host tree in C++17, evaluation in HIP kernels on the current GPU).  There is
no CPU fallback: FMM3D raises if the library or the GPU is unavailable.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsfmm.so")

POTENTIAL, CURL, POTGRAD = 0, 1, 2
FAR, NEAR, BOTH = 1, 2, 3
DIRECT_NOALLOC_MAX_TARGETS = 524288  # SFMM_DIRECT_NOALLOC_MAX_TARGETS
SELF_LIST = 4  # or-ed into `which`: the near field excludes the plan's self lists
FP64, MIXED = 0, 1
STAGES = ("gather", "P2M", "M2M", "M2L", "L2L", "L2P", "P2P", "scatter", "total")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int64)


class SfmmError(RuntimeError):
    pass


class TreeDesc(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("positions", _dp), ("max_level", ctypes.c_int),
                ("max_per_leaf", ctypes.c_int), ("vertical_axis", ctypes.c_int), ("phideg", ctypes.c_double),
                ("box_given", ctypes.c_int), ("corner", ctypes.c_double * 3), ("width", ctypes.c_double),
                ("threads", ctypes.c_int), ("on_device", ctypes.c_int)]


class TreeInfo(ctypes.Structure):
    _fields_ = [("depth", ctypes.c_int), ("n_node", ctypes.c_int64), ("n_leaf", ctypes.c_int64),
                ("m2l_pairs", ctypes.c_int64), ("near_pairs", ctypes.c_int64), ("width", ctypes.c_double),
                ("corner", ctypes.c_double * 3)]


class PlanDesc(ctypes.Structure):
    _fields_ = [("p", ctypes.c_int), ("precision", ctypes.c_int), ("max_ncharge", ctypes.c_int),
                ("self_row", _ip), ("self_val", _ip), ("m2l_kernel", ctypes.c_int)]


class BatchItem(ctypes.Structure):
    _fields_ = [("plan", ctypes.c_void_p), ("kind", ctypes.c_int), ("ncharge", ctypes.c_int), ("which", ctypes.c_int),
                ("dQ", ctypes.c_void_p), ("dU", ctypes.c_void_p)]


# symbol -> (restype, argtypes); every function declared in include/sfmm.h
SIGNATURES = {
    "sfmm_box_from_positions": (ctypes.c_int, [ctypes.c_int64, _dp, ctypes.c_int, ctypes.c_double, _dp]),
    "sfmm_tree_create": (ctypes.c_int, [ctypes.POINTER(TreeDesc), ctypes.POINTER(ctypes.c_void_p)]),
    "sfmm_tree_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "sfmm_tree_digest": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]),
    "sfmm_tree_get_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(TreeInfo)]),
    "sfmm_tree_level_counts": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _ip]),
    "sfmm_tree_level_lists": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int] + [_ip] * 11),
    "sfmm_tree_near_sources": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, _ip, ctypes.c_int64, _ip]),
    "sfmm_plan_slist": (ctypes.c_int, [ctypes.c_void_p, _ip, _ip, _ip]),
    "sfmm_plan_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(PlanDesc), ctypes.POINTER(ctypes.c_void_p)]),
    "sfmm_plan_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "sfmm_evaluate": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int]),
    "sfmm_evaluate_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_int]),
    "sfmm_plan_stream": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    "sfmm_plan_tune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
    "sfmm_plan_tiles": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "sfmm_plan_set_tiles": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    "sfmm_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "sfmm_last_stage_ms": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]),
    "sfmm_record_stages": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "sfmm_recorded_stage_ms": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_int)]),
    "sfmm_direct": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int64, _ip, _dp]),
    "sfmm_direct_device_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                                  ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]),
    "sfmm_batch_create": (ctypes.c_int, [ctypes.POINTER(BatchItem), ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_void_p)]),
    "sfmm_batch_launch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "sfmm_batch_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "sfmm_batch_streams": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "sfmm_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "sfmm_last_error": (ctypes.c_char_p, []),
}

_lib = None


def build(force: bool = False) -> None:
    """Compiles libsfmm.so in-tree (hipcc, gfx950)."""
    args = ["make", "-s", "-C", _HERE, "-j8"]
    if force:
        subprocess.run(["make", "-s", "-C", _HERE, "clean"], check=True)
    subprocess.run(args, check=True)


def lib():
    """Loads libsfmm.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SfmmError(f"{LIB_PATH} is missing: build it with `make -C {_HERE}` (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        raise SfmmError(f"sfmm error {rc}: {lib().sfmm_last_error().decode()}")


def _d(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _i(a: np.ndarray):
    return a.ctypes.data_as(_ip)


# ----------------------------------------------------------------- reference API
class Box:
    """Box(positions, vertical_axis=2, phideg=0) or Box(corner=..., width=...) (box.hpp:73-107)."""

    def __init__(self, positions=None, vertical_axis: int = 2, phideg: float = 0.0, *, corner=None, width=None):
        self.vertical_axis = int(vertical_axis)
        self.phideg = float(phideg)
        self.given = corner is not None
        if self.given:
            self.corner = np.asarray(corner, dtype=np.float64).reshape(3)
            self.width_ = float(width)
        elif positions is None:
            raise ValueError("Box needs positions or (corner, width)")


class OctreePartitioner:
    """OctreePartitioner(positions, box).refine(max_level, max_particles_per_node) (partition.hpp:380-446)."""

    def __init__(self, positions, box: Box):
        self._pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        self._box = box
        self.h = None

    def refine(self, max_level: int = 20, max_particles_per_node: int = 128, threads: int = 0,
               on_device: bool = False):
        """Builds the tree with `threads` host threads, or on the current GPU when
        `on_device` (the same tree, array for array)."""
        desc = TreeDesc()
        desc.n = self._pos.shape[0]
        desc.positions = _d(self._pos)
        desc.max_level = int(max_level)
        desc.max_per_leaf = int(max_particles_per_node)
        desc.vertical_axis = self._box.vertical_axis
        desc.phideg = self._box.phideg
        desc.box_given = int(self._box.given)
        if self._box.given:
            for d in range(3):
                desc.corner[d] = self._box.corner[d]
            desc.width = self._box.width_
        desc.threads = int(threads)
        desc.on_device = int(bool(on_device))
        h = ctypes.c_void_p()
        _check(lib().sfmm_tree_create(ctypes.byref(desc), ctypes.byref(h)))
        self.h = h
        return self

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.sfmm_tree_destroy(self.h)
            self.h = None

    @property
    def n(self) -> int:
        return self._pos.shape[0]

    def info(self) -> TreeInfo:
        info = TreeInfo()
        _check(lib().sfmm_tree_get_info(self.h, ctypes.byref(info)))
        return info

    def level(self) -> int:
        return self.info().depth

    def digest(self) -> int:
        """FNV-1a hash over every array of the tree; equal digests mean the same tree."""
        d = ctypes.c_uint64()
        _check(lib().sfmm_tree_digest(self.h, ctypes.byref(d)))
        return int(d.value)

    def near_sources(self, i: int) -> np.ndarray:
        """Original indices of the near-field sources of particle i (its leaf's near list)."""
        cnt = np.zeros(1, np.int64)
        _check(lib().sfmm_tree_near_sources(self.h, int(i), None, 0, _i(cnt)))
        out = np.zeros(int(cnt[0]), np.int64)
        _check(lib().sfmm_tree_near_sources(self.h, int(i), _i(out), len(out), _i(cnt)))
        return out

    def level_lists(self, l: int) -> dict:
        c = np.zeros(6, dtype=np.int64)
        _check(lib().sfmm_tree_level_counts(self.h, l, _i(c)))
        n_node, n_leaf, nidx, nil, ncl, nnl = (int(x) for x in c)
        depth = self.level()
        out = {
            "n_node": n_node, "n_leaf": n_leaf,
            "c_node": np.zeros(n_node, np.int64), "leaf2node": np.zeros(n_leaf, np.int64),
            "idx_row": np.zeros(n_leaf + 1, np.int64), "idx_val": np.zeros(nidx, np.int64),
            "ilist_row": np.zeros((n_node + 1) if l >= 2 else 1, np.int64), "ilist_val": np.zeros(4 * nil, np.int64),
            "clist_row": np.zeros((n_node + 1) if l < depth else 1, np.int64),
            "clist_val": np.zeros(2 * ncl, np.int64),
            "plist": np.zeros(2 * n_node if l >= 1 else 0, np.int64),
            "nlist_row": np.zeros(n_leaf + 1, np.int64), "nlist_val": np.zeros(2 * nnl, np.int64),
        }
        keys = ["c_node", "leaf2node", "idx_row", "idx_val", "ilist_row", "ilist_val", "clist_row", "clist_val",
                "plist", "nlist_row", "nlist_val"]
        _check(lib().sfmm_tree_level_lists(self.h, l, *[_i(out[k]) for k in keys]))
        return out


class FMM3D:
    """FMM3D(partitioner, p, ncharge) (fmm.hpp:134-882) on the current HIP device.

    rinv(Q, U[, gradient]) accumulates into U exactly as the reference; Q is
    (n, ncharge) (or (n,) when ncharge == 1), U is (n, ncharge) float64.
    rinv_potgrad(q, U) accumulates (phi, grad phi) of scalar charges, U (n, 4).
    """

    def __init__(self, partitioner: OctreePartitioner, p: int, ncharge: int = 3, precision: int = FP64,
                 m2l_kernel: int = 0, self_lists=None):
        """self_lists: optional (row, val) CSR of self_generator(i) (fmm.hpp:204),
        original particle indices; None means self_generator(i) = {i}."""
        if partitioner.h is None:
            raise SfmmError("refine() the partitioner before building an FMM3D")
        self.partitioner = partitioner  # the plan references the tree: keep it alive
        self.p = int(p)
        self.ncharge = int(ncharge)
        desc = PlanDesc()
        desc.p = self.p
        desc.precision = int(precision)
        desc.max_ncharge = max(1, min(4, self.ncharge))
        desc.m2l_kernel = int(m2l_kernel)
        if self_lists is not None:
            row = np.ascontiguousarray(self_lists[0], dtype=np.int64)
            val = np.ascontiguousarray(self_lists[1], dtype=np.int64)
            if row.shape != (partitioner.n + 1,):
                raise ValueError("self_lists row must have n + 1 entries")
            desc.self_row = _i(row)
            desc.self_val = _i(val)
        h = ctypes.c_void_p()
        _check(lib().sfmm_plan_create(partitioner.h, ctypes.byref(desc), ctypes.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.sfmm_plan_destroy(self.h)
            self.h = None

    @property
    def n(self) -> int:
        return self.partitioner.n

    def _eval(self, kind: int, ncharge: int, Q, U, which: int):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        if not (isinstance(U, np.ndarray) and U.dtype == np.float64 and U.flags.c_contiguous):
            raise TypeError("U must be a C-contiguous float64 numpy array (it is accumulated in place)")
        _check(lib().sfmm_evaluate(self.h, kind, ncharge, _d(Q), _d(U), which))
        return U

    def rinv(self, Q, U, gradient: bool = False, which: int = BOTH):
        """rinv<gradient>(Q, U): potentials sum Q_j/r, or (gradient) curl sum Q_j x R/|R|^3."""
        if gradient:
            return self._eval(CURL, 3, Q, U, which)
        return self._eval(POTENTIAL, self.ncharge, Q, U, which)

    def rinv_nonear(self, Q, U, gradient: bool = False):
        return self.rinv(Q, U, gradient, FAR)

    def N2N(self, Q, U, gradient: bool = False):
        return self.rinv(Q, U, gradient, NEAR)

    def rinv_potgrad(self, q, U, which: int = BOTH):
        """Scalar charges: U[:, 0] += phi, U[:, 1:4] += grad phi."""
        return self._eval(POTGRAD, 1, q, U, which)

    # device-resident path (torch tensors or raw pointers) ------------------
    def evaluate_device(self, kind: int, ncharge: int, q_ptr: int, u_ptr: int, which: int = BOTH):
        _check(lib().sfmm_evaluate_device(self.h, kind, ncharge, ctypes.c_void_p(q_ptr), ctypes.c_void_p(u_ptr),
                                          which))

    def synchronize(self):
        _check(lib().sfmm_synchronize(self.h))

    def tune(self, kind: int, ncharge: int, q_ptr: int):
        """Times the kernels' tile shapes with these device charges and keeps the fastest."""
        _check(lib().sfmm_plan_tune(self.h, kind, ncharge, ctypes.c_void_p(q_ptr)))

    def set_tiles(self, p2p_targets_per_lane: int, m2l_waves: int = 0):
        """Sets the tile shapes (near field 1, 2 or 4 targets per lane; dense M2L 4 or 8 wavefronts, 0 keeps)."""
        _check(lib().sfmm_plan_set_tiles(self.h, int(p2p_targets_per_lane), int(m2l_waves)))

    def tiles(self) -> dict:
        """{'p2p_targets_per_lane': T, 'm2l_waves': W} (W = 0 without the dense M2L)."""
        a, b = ctypes.c_int(), ctypes.c_int()
        _check(lib().sfmm_plan_tiles(self.h, ctypes.byref(a), ctypes.byref(b)))
        return {"p2p_targets_per_lane": a.value, "m2l_waves": b.value}

    def slist(self):
        """The far self pairs (_slist, fmm.hpp:398-408): (row, val) CSR over original indices."""
        nnz = np.zeros(1, np.int64)
        _check(lib().sfmm_plan_slist(self.h, None, None, _i(nnz)))
        row = np.zeros(self.n + 1, np.int64)
        val = np.zeros(max(int(nnz[0]), 1), np.int64)
        _check(lib().sfmm_plan_slist(self.h, _i(row), _i(val), _i(nnz)))
        return row, val[: int(nnz[0])]

    def stream(self) -> int:
        s = ctypes.c_void_p()
        _check(lib().sfmm_plan_stream(self.h, ctypes.byref(s)))
        return s.value or 0

    def stage_ms(self) -> dict:
        ms = (ctypes.c_float * 9)()
        _check(lib().sfmm_last_stage_ms(self.h, ms))
        return dict(zip(STAGES, [float(x) for x in ms]))

    def record_stages(self, count: int):
        """Keeps per-stage HIP-event times for each of the next `count` evaluations."""
        _check(lib().sfmm_record_stages(self.h, int(count)))

    def recorded_stage_ms(self, capacity: int) -> list:
        ms = (ctypes.c_float * (9 * max(capacity, 1)))()
        rec = ctypes.c_int()
        _check(lib().sfmm_recorded_stage_ms(self.h, ms, int(capacity), ctypes.byref(rec)))
        k = min(rec.value, capacity)
        return [dict(zip(STAGES, [float(ms[9 * r + s]) for s in range(9)])) for r in range(k)]

    def direct(self, kind: int, Q, targets, ncharge: int = 1):
        """FP64 direct sum at original-index targets (all sources j != i)."""
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        targets = np.ascontiguousarray(targets, dtype=np.int64)
        nout = {POTENTIAL: ncharge, CURL: 3, POTGRAD: 4}[kind]
        out = np.zeros((len(targets), nout))
        _check(lib().sfmm_direct(self.h, kind, ncharge, _d(Q), len(targets), _i(targets), _d(out)))
        return out

    def direct_device_noalloc(self, kind: int, ncharge: int, q_ptr: int, ntarget: int, t_ptr: int, out_ptr: int):
        """The direct sum with device arrays and no device allocation (static
        scratch; at most DIRECT_NOALLOC_MAX_TARGETS targets), asynchronous on the
        plan's stream: out[t, c] = value at original index targets[t]."""
        _check(lib().sfmm_direct_device_noalloc(self.h, kind, ncharge, ctypes.c_void_p(q_ptr), int(ntarget),
                                                ctypes.c_void_p(t_ptr), ctypes.c_void_p(out_ptr)))


class Batch:
    """Independent problems evaluated together over concurrent streams (sfmm_batch_*).

    items: sequence of (fmm, kind, ncharge, q_ptr, u_ptr[, which]) with device
    pointers (e.g. torch tensors' data_ptr()); the plans must be distinct and
    the arrays must outlive the batch.  streams=0 tunes the stream count.
    """

    def __init__(self, items, streams: int = 0):
        arr = (BatchItem * len(items))()
        self._plans = []
        for k, it in enumerate(items):
            fmm, kind, ncharge, qp, up = it[:5]
            which = it[5] if len(it) > 5 else BOTH
            arr[k] = BatchItem(fmm.h, kind, ncharge, which, ctypes.c_void_p(qp), ctypes.c_void_p(up))
            self._plans.append(fmm)  # keep the plans alive
        h = ctypes.c_void_p()
        _check(lib().sfmm_batch_create(arr, len(items), int(streams), ctypes.byref(h)))
        self.h = h

    @property
    def streams(self) -> int:
        v = ctypes.c_int()
        _check(lib().sfmm_batch_streams(self.h, ctypes.byref(v)))
        return v.value

    def launch(self, stream_ptr: int = 0):
        _check(lib().sfmm_batch_launch(self.h, ctypes.c_void_p(stream_ptr) if stream_ptr else None))

    def synchronize(self):
        _check(lib().sfmm_batch_synchronize(self.h))

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.sfmm_batch_destroy(self.h)
            self.h = None
