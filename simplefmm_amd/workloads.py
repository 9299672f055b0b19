"""Seeded synthetic particle sets for the FMM workloads.

All randomness comes from a counter-based splitmix64 stream, so the same seed
gives bit-identical inputs on any machine and in any language that implements
the same ten lines (the reference main.cpp:27-34 uses std::rand, which is not
portable; these generators replace it).

Workloads (sizes are the defaults; every generator takes n):
  uniform_cube   points uniform in [0,1)^3, charges uniform in [-1,1)
  plummer        Plummer sphere, scale radius 1, truncated at r=50, charges in (0,1]
  clusters       mixture of isotropic Gaussian clusters inside [0,1)^3
  sphere_shell   points on the unit sphere plus a fraction on an inner sphere,
                 standard-normal charges
"""
from __future__ import annotations

import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class Stream:
    """Counter-based splitmix64: value k of stream `seed` is mix(seed + (k+1)*golden)."""

    def __init__(self, seed: int):
        self.seed = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.count = 0

    def u64(self, n: int) -> np.ndarray:
        k = np.arange(self.count + 1, self.count + n + 1, dtype=np.uint64)
        self.count += n
        with np.errstate(over="ignore"):
            z = self.seed + k * _GOLDEN
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))

    def uniform(self, n: int) -> np.ndarray:
        """Doubles in [0,1) with 53 random bits."""
        return (self.u64(n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)

    def uniform_open0(self, n: int) -> np.ndarray:
        """Doubles in (0,1]."""
        return ((self.u64(n) >> np.uint64(11)).astype(np.float64) + 1.0) * (2.0 ** -53)

    def normal(self, n: int) -> np.ndarray:
        """Standard normals by the Box-Muller transform (cosine branch)."""
        u1 = self.uniform_open0(n)
        u2 = self.uniform(n)
        return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def uniform_cube(n: int, seed: int = 1, charge: str = "scalar"):
    """Positions uniform in [0,1)^3 (AoS, shape (n,3)); charges uniform in [-1,1)."""
    s = Stream(seed)
    pos = s.uniform(3 * n).reshape(n, 3)
    qdim = 3 if charge == "vector" else 1
    q = 2.0 * s.uniform(qdim * n) - 1.0
    return pos, (q.reshape(n, 3) if qdim == 3 else q)


def plummer(n: int, seed: int = 3, rmax: float = 50.0):
    """Plummer sphere with scale radius 1, truncated at rmax; charges uniform in (0,1].

    Radius by inverse transform of the enclosed mass M(r) = r^3/(1+r^2)^(3/2):
    u ~ U(0, M(rmax)), r = (u^(-2/3) - 1)^(-1/2).  Direction isotropic.
    """
    s = Stream(seed)
    mmax = rmax ** 3 / (1.0 + rmax * rmax) ** 1.5
    u = s.uniform_open0(n) * mmax
    r = 1.0 / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
    cos_t = 2.0 * s.uniform(n) - 1.0
    phi = 2.0 * math.pi * s.uniform(n)
    sin_t = np.sqrt(np.maximum(0.0, 1.0 - cos_t * cos_t))
    pos = np.stack([r * sin_t * np.cos(phi), r * sin_t * np.sin(phi), r * cos_t], axis=1)
    q = s.uniform_open0(n)
    return np.ascontiguousarray(pos), q


def clusters(n: int, seed: int = 5, kmax: int = 8, smin: float = 0.02, smax: float = 0.2):
    """1..kmax isotropic Gaussian clusters (sigma ~ U[smin,smax]) rejected to [0,1)^3."""
    s = Stream(seed)
    k = 1 + int(s.u64(1)[0] % np.uint64(kmax))
    centres = s.uniform(3 * k).reshape(k, 3)
    sig = smin + (smax - smin) * s.uniform(k)
    out = np.empty((n, 3))
    filled = 0
    while filled < n:
        m = 2 * (n - filled) + 64
        which = (s.u64(m) % np.uint64(k)).astype(np.int64)
        pts = centres[which] + sig[which, None] * s.normal(3 * m).reshape(m, 3)
        ok = np.all((pts >= 0.0) & (pts < 1.0), axis=1)
        take = pts[ok][: n - filled]
        out[filled: filled + len(take)] = take
        filled += len(take)
    q = 2.0 * s.uniform(n) - 1.0
    return out, q


def sphere_shell(n: int, seed: int = 7, inner_fraction: float = 0.1, inner_radius: float = 0.5):
    """(1-f) of the points on the unit sphere, f on a sphere of radius inner_radius.

    Directions are normalised Gaussians; charges are standard normal.
    """
    s = Stream(seed)
    g = s.normal(3 * n).reshape(n, 3)
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    n_in = int(round(inner_fraction * n))
    g[n - n_in:] *= inner_radius
    q = s.normal(n)
    return np.ascontiguousarray(g), q


def batch_specs(count: int = 256, seed: int = 11, nmin: int = 10_000, nmax: int = 1_000_000,
                orders=(6, 8, 10, 12)):
    """Sizes log-uniform in [nmin, nmax] and orders drawn from `orders` for a batch."""
    s = Stream(seed)
    u = s.uniform(count)
    sizes = np.rint(np.exp(math.log(nmin) + u * (math.log(nmax) - math.log(nmin)))).astype(np.int64)
    ps = np.array(orders)[(s.u64(count) % np.uint64(len(orders))).astype(np.int64)]
    return [(int(a), int(b), seed * 1000 + i) for i, (a, b) in enumerate(zip(sizes, ps))]


def write_harness_input(path: str, pos: np.ndarray, q: np.ndarray) -> None:
    """Binary input for the reference driver: 'SFMI', int32 1, int64 n, int32 qdim, pos, q."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64)
    n = pos.shape[0]
    qdim = 1 if q.ndim == 1 else q.shape[1]
    with open(path, "wb") as f:
        f.write(b"SFMI")
        f.write(np.int32(1).tobytes())
        f.write(np.int64(n).tobytes())
        f.write(np.int32(qdim).tobytes())
        f.write(pos.tobytes())
        f.write(q.tobytes())


def read_blocks(path: str) -> dict:
    """Reads an 'SFMO' output or 'SFMT' tree file of tagged blocks into a dict of arrays."""
    out = {}
    with open(path, "rb") as f:
        magic = f.read(4)
        if magic == b"SFMO":
            hdr = np.frombuffer(f.read(32), dtype=np.int64)
            out["header"] = hdr
            dtype = np.float64
        elif magic == b"SFMT":
            out["depth"] = int(np.frombuffer(f.read(8), dtype=np.int64)[0])
            dtype = np.int64
        else:
            raise ValueError(f"{path}: unknown magic {magic!r}")
        levels = []
        while True:
            tag = f.read(16)
            if len(tag) < 16:
                break
            name = tag.rstrip(b"\0").decode()
            count = int(np.frombuffer(f.read(8), dtype=np.int64)[0])
            arr = np.frombuffer(f.read(8 * count), dtype=dtype).copy()
            if magic == b"SFMT":
                if name == "head":
                    levels.append({})
                levels[-1][name] = arr
            else:
                out[name] = arr
        if magic == b"SFMT":
            out["levels"] = levels
    return out
