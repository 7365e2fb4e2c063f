"""Synthetic symmetric tridiagonal inputs and the ``.de.bin`` file format.

One PRNG for every workload: splitmix64 -> 53-bit uniforms in [0, 1), and
Box-Muller normals from consecutive uniform pairs (cos for the even member,
sin for the odd one).  Draw order within a matrix is fixed so that any
implementation of this spec reproduces the same bits:

* ``uniform``   D = 2u-1 (n draws), then E = 2u-1 (n-1 draws)
* ``normal``    D = N(0,1) (n normals), then E = N(0,1) (n-1 normals)
* ``toeplitz``  D = 2, E = 1 (no draws)
* ``clustered`` c_k = 2u-1 for k < 8, then D_i = c_(i mod 8) + 1e-8 N(0,1),
                then E_i = 1e-6 u

Seeds: ``0xB7DC0000 + 1000*config + n (+ matrix index when batched)``.

The ``.de.bin`` layout is the reference harness's compact tridiagonal file:
int64 n, then n FP64 diagonal entries, then n-1 FP64 off-diagonal entries
(gpu/bench/tridiag_dc_bench.cu:393-463 reads it; a dense n*n FP64 payload
after the header is also accepted there and here).
"""
from __future__ import annotations

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
SEED_BASE = 0xB7DC0000
GENERATORS = ("uniform", "normal", "toeplitz", "clustered")


def seed_for(config: int, n: int, index: int = 0) -> int:
    return SEED_BASE + 1000 * int(config) + int(n) + int(index)


class SplitMix64:
    """Counter-based splitmix64: output k is mix(seed + (k+1)*gamma)."""

    def __init__(self, seed: int):
        self.seed = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.count = 0

    def raw(self, k: int) -> np.ndarray:
        idx = np.arange(self.count + 1, self.count + 1 + k, dtype=np.uint64)
        self.count += k
        with np.errstate(over="ignore"):
            z = self.seed + idx * _GAMMA
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, k: int) -> np.ndarray:
        """k doubles in [0, 1) with 53 random bits."""
        return (self.raw(k) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)

    def normal(self, k: int) -> np.ndarray:
        m = (k + 1) // 2
        u = self.uniform(2 * m)
        u1, u2 = u[0::2], u[1::2]
        r = np.sqrt(-2.0 * np.log1p(-u1))
        out = np.empty(2 * m, dtype=np.float64)
        out[0::2] = r * np.cos(2.0 * np.pi * u2)
        out[1::2] = r * np.sin(2.0 * np.pi * u2)
        return out[:k]


def generate(kind: str, n: int, seed: int) -> tuple[np.ndarray, np.ndarray]:
    """Return (D, E) for one n x n symmetric tridiagonal of the given kind."""
    if n < 0:
        raise ValueError("n must be >= 0")
    g = SplitMix64(seed)
    ne = max(n - 1, 0)
    if kind == "uniform":
        d = 2.0 * g.uniform(n) - 1.0
        e = 2.0 * g.uniform(ne) - 1.0
    elif kind == "normal":
        d = g.normal(n)
        e = g.normal(ne)
    elif kind == "toeplitz":
        d = np.full(n, 2.0)
        e = np.full(ne, 1.0)
    elif kind == "clustered":
        c = 2.0 * g.uniform(8) - 1.0
        d = c[np.arange(n) % 8] + 1e-8 * g.normal(n)
        e = 1e-6 * g.uniform(ne)
    else:
        raise ValueError(f"unknown generator {kind!r}; expected one of {GENERATORS}")
    return np.ascontiguousarray(d), np.ascontiguousarray(e)


def batched_sizes(count: int, lo: int, hi: int, seed: int) -> np.ndarray:
    """Sizes round(exp(U(ln lo, ln hi))) for the batched workload."""
    g = SplitMix64(seed)
    u = g.uniform(count)
    s = np.rint(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))).astype(np.int64)
    return np.clip(s, lo, hi)


def batched_workload(count: int = 2048, lo: int = 32, hi: int = 4096, config: int = 4,
                     kind: str = "normal"):
    """The batched workload: ``count`` sizes log-uniform in [lo, hi] (sizes
    seed ``seed_for(config, 0)``), matrix b drawn by ``generate(kind, n_b,
    seed_for(config, n_b, b))``.  Returns (sizes, [(D, E), ...])."""
    sizes = batched_sizes(count, lo, hi, seed_for(config, 0))
    mats = [generate(kind, int(nb), seed_for(config, int(nb), b)) for b, nb in enumerate(sizes)]
    return sizes, mats


def pack_batch(mats) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Concatenate matrices for the batched entry: (D rows, E rows, offsets).
    Matrix b owns rows [off_b, off_b+n_b); its E sits in the first n_b - 1
    of them and the last row's E slot is 0."""
    sizes = np.array([d.shape[0] for d, _ in mats], dtype=np.int64)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    rows = int(off[-1])
    d_cat = np.zeros(max(rows, 1))
    e_cat = np.zeros(max(rows, 1))
    for b, (d, e) in enumerate(mats):
        o, nb = int(off[b]), int(sizes[b])
        d_cat[o:o + nb] = d
        e_cat[o:o + max(nb - 1, 0)] = e[:max(nb - 1, 0)]
    return d_cat[:rows], e_cat[:rows], off


def write_de_bin(path: str, d: np.ndarray, e: np.ndarray) -> None:
    n = int(d.shape[0])
    if e.shape[0] != max(n - 1, 0):
        raise ValueError("E must have n-1 entries")
    with open(path, "wb") as f:
        f.write(np.int64(n).tobytes())
        f.write(np.ascontiguousarray(d, dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(e, dtype=np.float64).tobytes())


def read_de_bin(path: str) -> tuple[np.ndarray, np.ndarray]:
    raw = np.fromfile(path, dtype=np.uint8)
    if raw.size < 8:
        raise ValueError(f"{path}: missing int64 header")
    n = int(raw[:8].view(np.int64)[0])
    if n <= 0:
        raise ValueError(f"{path}: invalid n={n}")
    payload = raw[8:]
    if payload.size == (2 * n - 1) * 8:
        vals = payload.view(np.float64)
        return vals[:n].copy(), vals[n:].copy()
    if payload.size == n * n * 8:
        a = payload.view(np.float64).reshape(n, n)  # column-major dense
        d = np.diagonal(a).copy()
        e = np.array([a[i, i + 1] for i in range(n - 1)], dtype=np.float64)
        return d, e
    raise ValueError(f"{path}: payload of {payload.size} bytes fits neither "
                     f"compact (2n-1) nor dense (n^2) FP64 layout for n={n}")
